// Host engine of the gfx950 modular-squaring path and its C ABI.
//
// Replaces the reference's OpenCL engine: transformGPU<LN> (src/transformGPU.h),
// engine<LN> (src/engine.h) and the OpenCL wrapper (src/ocl.h).  One context
// owns one modulus N, its transform plan, twiddle tables, pre-transformed
// constants and the register file, all resident in HBM, and issues the
// kernels of kernels.hip on its own HIP stream.
#include <gmp.h>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/annie_hip.h"
#include "kernels.h"
#include "plan.h"

namespace annie {

template <int M, int MODE>
__global__ void k_row(const NttArgs A);
template <int M>
__global__ void k_col_fwd(const NttArgs A);
template <int M>
__global__ void k_col_inv(const NttArgs A);
__global__ void k_scale(uint32_t* x, size_t n, int K, NttArgs A, uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3,
                        uint32_t s4);
template <int K, int QS>
__global__ void k_carry(const CarryArgs A);
template <int K>
__global__ void k_digits_to_res(const uint32_t* digits, uint32_t* dst, const CarryArgs A);
__global__ void k_copy_slots(uint32_t* base, size_t dst_off, size_t src_off, size_t len, const uint32_t* slot_ids,
                             size_t slot_stride);
template <int M, int K, int QS>
__global__ void k_sqr_run(const RunArgs R);

static thread_local std::string g_last_error;

static int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIPCHK(expr)                                                                               \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return fail(ANNIE_ERR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Register file layout: every register is K planes of n residues.
enum Reg : int {
  R0 = 0, R1 = 1, R2 = 2,  // user registers
  RONE,                    // MMGW form of 1
  RTR,                     // T_R = a x^2 mod R, then m' = T_R N^-1 mod R
  RWR, RWQ,                // transform work buffers (cyclic / negacyclic)
  RSR, RSQ,                // forward transforms of a runtime operand (mul)
  RNQ,                     // forward negacyclic transform of N, scaled by n^-1 2^32
  RNIR,                    // forward cyclic transform of N^-1 mod R, scaled by n^-1 2^32
  RTMP,                    // scratch for set-up uploads
  RSNAP,                   // batches: copy of r0 taken by a flagged square_mul
  NREG
};

struct Mpz {
  mpz_t v;
  Mpz() { mpz_init(v); }
  ~Mpz() { mpz_clear(v); }
  Mpz(const Mpz&) = delete;
  Mpz& operator=(const Mpz&) = delete;
};

// Ring of pinned staging buffers for the per-operation slot lists and
// multipliers of a batch (entry e is rewritten only after its previous
// host-to-device copy has completed).
constexpr int NSTAGE = 16;
struct Stage {
  uint32_t* dev = nullptr;
  uint32_t* host = nullptr;
  size_t cap = 0;  // words per entry
  int pos = 0;
  hipEvent_t ev[NSTAGE] = {};
  bool pending[NSTAGE] = {};
};

struct Ctx {
  Plan pl;
  int device = 0;
  int m1 = 0, m2 = 0;
  // Launch schedule of the NTT passes (0: the default): lines per row-pass
  // workgroup for rows of fewer than 64 threads, columns per column-pass
  // workgroup for columns of up to 2^9 (annie_hip_set_schedule).
  uint32_t sched_lines = 0, sched_cols = 0;
  size_t n = 0;
  hipStream_t stream = nullptr;
  // Stream of the next kernel launch: `stream`, or `aux` while the
  // negacyclic half of a squaring is issued (single two-pass contexts).
  hipStream_t cur = nullptr;
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool dual_stream = true;
  // One allocation of nslots_alloc slots; a slot is NREG registers of K * n
  // words followed by 2 * DIGIT_LIMBS * n words of raw digits (T_Q, get()).
  uint32_t* mem = nullptr;
  size_t nslots_alloc = 1, slot_stride = 0;
  uint32_t* regs = nullptr;   // slot 0 registers
  uint32_t* raw = nullptr;    // slot 0 raw digits
  // Slot list of the operation in progress (device array, null: slot 0 only).
  const uint32_t* slot_ids = nullptr;
  int nslots = 1;
  Stage stage;
  Stage run_stage;  // slot lists and multipliers of squaring runs (allocated on first use)
  std::vector<std::unique_ptr<Mpz>> slotN;  // modulus of each slot (batches)
  // Bits per coefficient of each slot (batches: moduli of one (K, LN) whose
  // plans differ in lbase share a batch), and their device copy.
  std::vector<int> slot_cb;
  // One device allocation of 3 words per slot: digit widths (cbits_dev),
  // bases (base_dev: the test's a, for multiplier bits) and the slot list of
  // the last batched squaring with short-list bits (list_dev, whose host copy
  // is list_host; re-uploaded only when the list changes).
  uint32_t* cbits_dev = nullptr;
  uint32_t* base_dev = nullptr;
  uint32_t* list_dev = nullptr;
  std::vector<uint32_t> slot_base, list_host;
  // Per-position bits of the batched squaring being issued (bits_count
  // positions, 0 = none): multipliers for its first carry, snapshot flags for
  // its last.
  size_t bits_count = 0;
  uint32_t mulbits[INLINE_SLOTS / 32] = {}, snapbits[INLINE_SLOTS / 32] = {};
  // Batched squaring replayed from captured HIP graphs: the slot list and
  // multipliers live at fixed device addresses (graph_buf: ids, then
  // multipliers), one executable graph per list length (grid sizes).
  bool use_graphs = false;
  uint32_t* graph_buf = nullptr;
  struct Graph {
    size_t count;
    hipGraphExec_t exec;
    uint64_t used;
  };
  std::vector<Graph> graphs;
  uint64_t graph_clock = 0;
  TwDev* tw = nullptr;        // 4 tables * K * n roots (Montgomery form)
  mpz_t N, factor;
  bool init_failed = false;
  PrimeConsts pc;
  NttArgs nbase{};
  CarryArgs cbase{};
  // Per-kernel event timing (diagnostics / benchmark roofline).
  bool profiling = false;
  std::vector<hipEvent_t> ev;        // pool, two per record
  std::vector<int> rec_kind;         // kernel kind of each record
  std::vector<std::string> kinds;    // kind names

  uint32_t* reg(int r) const { return regs + size_t(r) * size_t(pl.K) * n; }
  uint32_t* tq() const { return raw; }
  uint32_t* rawtmp() const { return raw + size_t(DIGIT_LIMBS) * n; }
  const TwDev* fwdR() const { return tw + 0 * size_t(pl.K) * n; }
  const TwDev* invR() const { return tw + 1 * size_t(pl.K) * n; }
  const TwDev* fwdQ() const { return tw + 2 * size_t(pl.K) * n; }
  const TwDev* invQ() const { return tw + 3 * size_t(pl.K) * n; }
};

}  // namespace annie

struct annie_hip {
  annie::Ctx c;
};

namespace annie {

// Bits per coefficient of `slot` (the plan's, without per-slot values).
static int slot_cbits(const Ctx& c, size_t slot) { return c.slot_cb.empty() ? c.pl.cbits() : c.slot_cb[slot]; }

static NttJob job(const uint32_t* src, uint32_t* dst, const TwDev* f, const TwDev* i, const uint32_t* op = nullptr) {
  NttJob j;
  j.src = src;
  j.dst = dst;
  j.fwd = f;
  j.inv = i;
  j.opnd = op;
  return j;
}

// Records an event pair around the next launch when profiling is on.
struct ProfScope {
  Ctx& c;
  size_t rec = size_t(-1);
  ProfScope(Ctx& cc, const std::string& name) : c(cc) {
    if (!c.profiling) return;
    int kind = -1;
    for (size_t i = 0; i < c.kinds.size(); ++i)
      if (c.kinds[i] == name) kind = int(i);
    if (kind < 0) {
      kind = int(c.kinds.size());
      c.kinds.push_back(name);
    }
    rec = c.rec_kind.size();
    while (c.ev.size() < 2 * (rec + 1)) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) {
        rec = size_t(-1);
        return;
      }
      c.ev.push_back(e);
    }
    c.rec_kind.push_back(kind);
    (void)hipEventRecord(c.ev[2 * rec], c.cur);
  }
  ~ProfScope() {
    if (rec != size_t(-1)) (void)hipEventRecord(c.ev[2 * rec + 1], c.cur);
  }
};

static const char* row_mode_name(int mode) {
  return mode == ROW_FWD ? "fwd" : mode == ROW_SQR ? "sqr" : mode == ROW_MULC ? "mulc" : "mulv";
}

template <int M>
static void launch_row_m(int mode, dim3 g, dim3 b, size_t shm, hipStream_t s, const NttArgs& A) {
  switch (mode) {
    case ROW_FWD: hipLaunchKernelGGL((k_row<M, ROW_FWD>), g, b, shm, s, A); break;
    case ROW_SQR: hipLaunchKernelGGL((k_row<M, ROW_SQR>), g, b, shm, s, A); break;
    case ROW_MULC: hipLaunchKernelGGL((k_row<M, ROW_MULC>), g, b, shm, s, A); break;
    default: hipLaunchKernelGGL((k_row<M, ROW_MULV>), g, b, shm, s, A); break;
  }
}

template <int M>
static void launch_col_m(bool inverse, dim3 g, dim3 b, size_t shm, hipStream_t s, const NttArgs& A) {
  if (inverse) hipLaunchKernelGGL((k_col_inv<M>), g, b, shm, s, A);
  else hipLaunchKernelGGL((k_col_fwd<M>), g, b, shm, s, A);
}

// Default launch geometry of the passes.  Row pass: one 64-lane wave per
// workgroup for short rows (64 / Tl lines of Tl threads), one line per
// workgroup otherwise.  Column pass: columns of 2^10 and 2^11 take 8 columns
// per workgroup (a compile-time constant of their kernels); shorter columns
// at most 256 threads and 16 columns per workgroup, and at most the number
// of columns (columns of 2^9: 8 rather than 16, measured 3% faster per
// squaring at 2^19 by the schedule autotuner and even at 2^18).
static uint32_t row_lines_default(int m1, int m2) {
  const uint32_t Tl = (1u << m2) / NTT_R, rows = 1u << m1;
  uint32_t lines = (Tl >= 64) ? 1u : 64u / Tl;
  return lines > rows ? rows : lines;
}

static uint32_t col_cols_default(int m1, int LN) {
  if (m1 >= 10) return 8u;
  const uint32_t Tl = (1u << m1) / NTT_R, n2 = 1u << (LN - m1);
  uint32_t C = 256u / Tl;
  if (C > 16) C = 16;
  if (C > n2) C = n2;
  return C;
}

// Whether a schedule fits the passes of a plan with column levels m1 and row
// levels m2 (0 = default for either).  Row lines: a power of two, at most the
// number of rows, at most 512 threads and 64 KiB of LDS per workgroup, rows
// of fewer than 64 threads only.  Column count: a power of two, at most the
// number of columns, at most 1024 threads and 64 KiB of LDS (data and staged
// roots) per workgroup, columns of up to 2^9 only.
static bool schedule_fits(int m1, int m2, int LN, uint32_t lines, uint32_t cols) {
  auto pow2 = [](uint32_t v) { return v != 0 && (v & (v - 1)) == 0; };
  if (lines != 0) {
    const uint32_t S = 1u << m2, Tl = S / NTT_R, rows = 1u << m1;
    if (!pow2(lines) || Tl >= 64 || lines > rows || lines * Tl > 512u) return false;
    if (size_t(lines) * (S + (S >> 5)) * 4 > 65536) return false;
  }
  if (cols != 0) {
    if (m1 == 0 || m1 >= 10) return false;
    const uint32_t S = 1u << m1, Tl = S / NTT_R, n2 = 1u << (LN - m1);
    if (!pow2(cols) || cols > n2 || cols * Tl > 1024u) return false;
    if (size_t(cols + 1) * S * 4 > 65536) return false;
  }
  return true;
}

static int launch_row(Ctx& c, int mode, const NttJob* jobs, int nj) {
  NttArgs A = c.nbase;
  for (int i = 0; i < nj; ++i) A.job[i] = jobs[i];
  A.slot_ids = c.slot_ids;
  A.slot_stride = c.slot_stride;
  A.njobs = nj;
  A.m = c.m2;
  A.m1 = c.m1;
  const uint32_t S = 1u << c.m2, Tl = S / NTT_R;
  const uint32_t rows = 1u << c.m1;
  uint32_t lines = row_lines_default(c.m1, c.m2);
  if (c.sched_lines != 0 && Tl < 64) lines = c.sched_lines;  // checked by annie_hip_set_schedule
  A.cols = int(lines);
  const dim3 grid(rows / lines, c.pl.K, nj * c.nslots), block(lines * Tl);
  const size_t shm = size_t(lines) * (S + (S >> 5)) * sizeof(uint32_t);
  ProfScope ps(c, "k_row<" + std::to_string(c.m2) + "," + row_mode_name(mode) + ">x" + std::to_string(nj));
  switch (c.m2) {
    case 4: launch_row_m<4>(mode, grid, block, shm, c.cur, A); break;
    case 5: launch_row_m<5>(mode, grid, block, shm, c.cur, A); break;
    case 6: launch_row_m<6>(mode, grid, block, shm, c.cur, A); break;
    case 7: launch_row_m<7>(mode, grid, block, shm, c.cur, A); break;
    case 8: launch_row_m<8>(mode, grid, block, shm, c.cur, A); break;
    case 9: launch_row_m<9>(mode, grid, block, shm, c.cur, A); break;
    case 10: launch_row_m<10>(mode, grid, block, shm, c.cur, A); break;
    case 11: launch_row_m<11>(mode, grid, block, shm, c.cur, A); break;
    case 12: launch_row_m<12>(mode, grid, block, shm, c.cur, A); break;
    case 13: launch_row_m<13>(mode, grid, block, shm, c.cur, A); break;
    default: return fail(ANNIE_ERR_SIZE, "no row kernel for this length");
  }
  HIPCHK(hipGetLastError());
  return ANNIE_OK;
}

static int launch_col(Ctx& c, bool inverse, const NttJob* jobs, int nj) {
  NttArgs A = c.nbase;
  for (int i = 0; i < nj; ++i) A.job[i] = jobs[i];
  A.slot_ids = c.slot_ids;
  A.slot_stride = c.slot_stride;
  A.njobs = nj;
  A.m = c.m1;
  A.m1 = c.m1;
  // Columns of 2^10 and 2^11: 32 elements per thread, 8 columns per block,
  // roots from the global table.  Shorter columns: 16 elements per thread, up
  // to 16 columns per block, roots staged in LDS after the column data.
  const bool wide = c.m1 >= 10;
  const uint32_t S = 1u << c.m1, Tl = S / (wide ? 32u : uint32_t(NTT_R));
  const uint32_t n2 = uint32_t(c.n >> c.m1);
  uint32_t C = col_cols_default(c.m1, c.pl.LN);
  if (c.sched_cols != 0 && !wide) C = c.sched_cols;  // checked by annie_hip_set_schedule
  // The 32-element column kernels take 8 columns per block as a compile-time
  // constant (n2 >= 2^10 whenever m1 >= 10, so C is never clipped there).
  if (wide && C != 8u) return fail(ANNIE_ERR_SIZE, "column pass geometry");
  A.cols = int(C);
  const dim3 grid(n2 / C, c.pl.K, nj * c.nslots), block(C * Tl);
  const size_t shm = size_t(wide ? C : C + 1) * S * sizeof(uint32_t);
  ProfScope ps(c, std::string(inverse ? "k_col_inv<" : "k_col_fwd<") + std::to_string(c.m1) + ">x" + std::to_string(nj));
  switch (c.m1) {
    case 7: launch_col_m<7>(inverse, grid, block, shm, c.cur, A); break;
    case 8: launch_col_m<8>(inverse, grid, block, shm, c.cur, A); break;
    case 9: launch_col_m<9>(inverse, grid, block, shm, c.cur, A); break;
    case 10: launch_col_m<10>(inverse, grid, block, shm, c.cur, A); break;
    case 11: launch_col_m<11>(inverse, grid, block, shm, c.cur, A); break;
    default: return fail(ANNIE_ERR_SIZE, "no column kernel for this length");
  }
  HIPCHK(hipGetLastError());
  return ANNIE_OK;
}

// Forward transform (no pointwise): src -> dst for each job.
static int forward(Ctx& c, const NttJob* jobs, int nj) {
  int rc;
  if (c.m1 > 0) {
    if ((rc = launch_col(c, false, jobs, nj))) return rc;
    NttJob j2[2];
    for (int i = 0; i < nj; ++i) {
      j2[i] = jobs[i];
      j2[i].src = jobs[i].dst;
    }
    return launch_row(c, ROW_FWD, j2, nj);
  }
  return launch_row(c, ROW_FWD, jobs, nj);
}

// Convolution: dst = inverse(pointwise(forward(src))), for each job.
static int convolve(Ctx& c, int mode, const NttJob* jobs, int nj) {
  int rc;
  if (c.m1 > 0) {
    if ((rc = launch_col(c, false, jobs, nj))) return rc;
    NttJob j2[2];
    for (int i = 0; i < nj; ++i) {
      j2[i] = jobs[i];
      j2[i].src = jobs[i].dst;
    }
    if ((rc = launch_row(c, mode, j2, nj))) return rc;
    return launch_col(c, true, j2, nj);
  }
  return launch_row(c, mode, jobs, nj);
}

// bits (batched squarings with short-list bits, c.bits_count > 0): 1 = this
// carry takes the per-position multipliers, 2 = the snapshot flags.
static int launch_carry(Ctx& c, const CarryJob* jobs, int nj, const uint32_t* a_slot = nullptr,
                        const uint32_t* snap = nullptr, int bits = 0) {
  CarryArgs A = c.cbase;
  for (int i = 0; i < nj; ++i) A.job[i] = jobs[i];
  A.slot_ids = c.slot_ids;
  A.slot_stride = c.slot_stride;
  A.njobs = nj;
  A.a_slot = a_slot;
  A.snap = snap;
  if (bits != 0 && c.bits_count > 0) {
    A.bits_mode = bits;
    A.slot_base = c.base_dev;
    std::memcpy(A.bits, bits == 1 ? c.mulbits : c.snapbits, sizeof(A.bits));
  }
  const size_t nb = c.n / CARRY_L;
  // CARRY_WG - 1 owned blocks per workgroup (thread 0 only hands over a carry).
  const dim3 grid(unsigned((nb + CARRY_WG - 2) / (CARRY_WG - 1)), nj * c.nslots), block(CARRY_WG);
  const int qs = c.pl.cbits() >> 5;
  if (qs > 2) return fail(ANNIE_ERR_SIZE, "digit width above 95 bits");
  ProfScope ps(c, "k_carry<" + std::to_string(c.pl.K) + ">x" + std::to_string(nj));
  switch (c.pl.K * 4 + qs) {
#define ANNIE_CARRY_CASE(K, QS) \
  case K * 4 + QS: hipLaunchKernelGGL((k_carry<K, QS>), grid, block, 0, c.cur, A); break;
    ANNIE_CARRY_CASE(2, 0)
    ANNIE_CARRY_CASE(2, 1)
    ANNIE_CARRY_CASE(2, 2)
    ANNIE_CARRY_CASE(3, 0)
    ANNIE_CARRY_CASE(3, 1)
    ANNIE_CARRY_CASE(3, 2)
    ANNIE_CARRY_CASE(4, 0)
    ANNIE_CARRY_CASE(4, 1)
    ANNIE_CARRY_CASE(4, 2)
    ANNIE_CARRY_CASE(5, 0)
    ANNIE_CARRY_CASE(5, 1)
    ANNIE_CARRY_CASE(5, 2)
#undef ANNIE_CARRY_CASE
    default: return fail(ANNIE_ERR_SIZE, "no carry kernel for this RNS size");
  }
  HIPCHK(hipGetLastError());
  return ANNIE_OK;
}

static CarryJob cjob(const uint32_t* src, uint32_t* dst, uint32_t a, int nega, int canon, int raw_out,
                     const uint32_t* sub = nullptr) {
  CarryJob j{};
  j.src = src;
  j.dst = dst;
  j.sub = sub;
  j.a = a;
  j.nega = nega;
  j.canon = canon;
  j.raw_out = raw_out;
  return j;
}

// Buffers of one squaring: the operand x, the result y (may equal x), and the
// work registers: wr / wq the cyclic / negacyclic convolutions, tr T_R then
// m' + 2, sq the (m' + 2) N product, tq the T_Q digits (DIGIT_LIMBS planes).
struct SqBufs {
  const uint32_t* x;
  uint32_t* y;
  uint32_t *wr, *wq, *tr, *sq, *tq;
};

// The context's own work registers, r0 in and out.
static SqBufs ctx_bufs(const Ctx& c) {
  return SqBufs{c.reg(R0), c.reg(R0), c.reg(RWR), c.reg(RWQ), c.reg(RTR), c.reg(RSQ), c.tq()};
}

// reduceMMGW (src/transformGPU.h:523-538): with T_R in B.tr and T_Q (raw
// digits) in B.tq: m' = T_R N^-1 mod R, canonical, + 2; dst = (m' + 2) N - T_Q
// mod Q.  snap (batches): per-slot flags; a flagged slot also gets the result
// in RSNAP.  join: T_Q is being produced on the aux stream; wait for it before
// the last carry.  The (m' + 2) N product works in B.sq, so B.wq may be in
// use on aux.
static int reduce(Ctx& c, const SqBufs& B, uint32_t* dst, bool raw_out, const uint32_t* snap = nullptr,
                  bool join = false, bool bits = false) {
  int rc;
  NttJob j = job(B.tr, B.wr, c.fwdR(), c.invR(), c.reg(RNIR));
  if ((rc = convolve(c, ROW_MULC, &j, 1))) return rc;
  CarryJob cj = cjob(B.wr, B.tr, 1u, 0, 1, 0);
  if ((rc = launch_carry(c, &cj, 1))) return rc;
  j = job(B.tr, B.sq, c.fwdQ(), c.invQ(), c.reg(RNQ));
  if ((rc = convolve(c, ROW_MULC, &j, 1))) return rc;
  if (join) HIPCHK(hipStreamWaitEvent(c.stream, c.ev_join, 0));
  cj = cjob(B.sq, dst, 1u, 1, 0, raw_out ? 1 : 0, B.tq);
  if (snap != nullptr || bits) cj.dst2 = c.reg(RSNAP);
  return launch_carry(c, &cj, 1, nullptr, snap, bits ? 2 : 0);
}

static int reduce(Ctx& c, uint32_t* dst, bool raw_out, const uint32_t* snap = nullptr, bool join = false) {
  return reduce(c, ctx_bufs(c), dst, raw_out, snap, join);
}

// squareMMGW (src/transformGPU.h:556-568): B.y = a B.x^2.  a_slot: per-slot
// multipliers of a batch operation (device array in slot-list order),
// replacing a.
static int square_mul(Ctx& c, const SqBufs& B, uint32_t a, const uint32_t* a_slot = nullptr, bool bits = false) {
  int rc;
  if (c.aux != nullptr && a_slot == nullptr && !bits && !c.profiling) {
    // Two streams (not while per-kernel timing is on, so that each kernel is
    // timed alone): the negacyclic half (x^2 mod Q -> T_Q) on aux runs beside
    // the m' chain on the main stream, whose one-job passes leave much of the
    // GPU idle; it is forked after the cyclic half's carry (0.5% faster than
    // forking before the cyclic half) and the last carry joins them.
    const NttJob jr = job(B.x, B.wr, c.fwdR(), c.invR());
    if ((rc = convolve(c, ROW_SQR, &jr, 1))) return rc;
    const CarryJob cr = cjob(B.wr, B.tr, a, 0, 0, 0);
    if ((rc = launch_carry(c, &cr, 1))) return rc;
    HIPCHK(hipEventRecord(c.ev_fork, c.stream));
    HIPCHK(hipStreamWaitEvent(c.aux, c.ev_fork, 0));
    c.cur = c.aux;
    const NttJob jq = job(B.x, B.wq, c.fwdQ(), c.invQ());
    rc = convolve(c, ROW_SQR, &jq, 1);
    if (rc == ANNIE_OK) {
      const CarryJob cq = cjob(B.wq, B.tq, a, 1, 0, 1);
      rc = launch_carry(c, &cq, 1);
    }
    c.cur = c.stream;
    if (rc) return rc;
    HIPCHK(hipEventRecord(c.ev_join, c.aux));
    return reduce(c, B, B.y, false, nullptr, true);
  }
  const NttJob j[2] = {job(B.x, B.wr, c.fwdR(), c.invR()), job(B.x, B.wq, c.fwdQ(), c.invQ())};
  if ((rc = convolve(c, ROW_SQR, j, 2))) return rc;
  const CarryJob cj[2] = {cjob(B.wr, B.tr, a, 0, 0, 0), cjob(B.wq, B.tq, a, 1, 0, 1)};
  if ((rc = launch_carry(c, cj, 2, a_slot, nullptr, bits ? 1 : 0))) return rc;
  return reduce(c, B, B.y, false, a_slot, false, bits);
}

static int square_mul(Ctx& c, uint32_t a, const uint32_t* a_slot = nullptr, bool bits = false) {
  return square_mul(c, ctx_bufs(c), a, a_slot, bits);
}

// mulMMGW (src/transformGPU.h:540-554).
static int mul(Ctx& c, int dst, int src) {
  int rc;
  const NttJob f[2] = {job(c.reg(src), c.reg(RSR), c.fwdR(), c.invR()), job(c.reg(src), c.reg(RSQ), c.fwdQ(), c.invQ())};
  if ((rc = forward(c, f, 2))) return rc;
  const NttJob j[2] = {job(c.reg(dst), c.reg(RWR), c.fwdR(), c.invR(), c.reg(RSR)),
                       job(c.reg(dst), c.reg(RWQ), c.fwdQ(), c.invQ(), c.reg(RSQ))};
  if ((rc = convolve(c, ROW_MULV, j, 2))) return rc;
  const CarryJob cj[2] = {cjob(c.reg(RWR), c.reg(RTR), 1u, 0, 0, 0), cjob(c.reg(RWQ), c.tq(), 1u, 1, 0, 1)};
  if ((rc = launch_carry(c, cj, 2))) return rc;
  return reduce(c, c.reg(dst), false);
}

static int copy_reg(Ctx& c, int dst, int src) {
  if (dst == src) return ANNIE_OK;
  const size_t len = size_t(c.pl.K) * c.n;
  if (c.slot_ids == nullptr) {
    HIPCHK(hipMemcpyAsync(c.reg(dst), c.reg(src), len * 4, hipMemcpyDeviceToDevice, c.stream));
    return ANNIE_OK;
  }
  const unsigned blocks = unsigned(std::min<size_t>((len / 4 + 255) / 256, 64));
  hipLaunchKernelGGL(k_copy_slots, dim3(blocks, c.nslots), dim3(256), 0, c.stream, c.mem, size_t(dst) * len,
                     size_t(src) * len, len, c.slot_ids, c.slot_stride);
  HIPCHK(hipGetLastError());
  return ANNIE_OK;
}

// Signed digits (DIGIT_LIMBS planes, two's complement) -> integer sum d_i B^i.
static void digits_to_mpz(const Ctx& c, const std::vector<uint32_t>& d, int cb, mpz_t z) {
  const size_t n = c.n;
  const size_t nbits = n * size_t(cb);
  const size_t nl = nbits / 64 + 3;
  std::vector<uint64_t> limbs(nl, 0);
  __int128 carry = 0;
  size_t pos = 0;
  const unsigned __int128 mask = (cb >= 128) ? ~(unsigned __int128)0 : (((unsigned __int128)1 << cb) - 1);
  for (size_t i = 0; i < n; ++i) {
    const unsigned __int128 u =
        (unsigned __int128)d[i] | ((unsigned __int128)d[n + i] << 32) | ((unsigned __int128)d[2 * n + i] << 64);
    // sign-extend the 96-bit digit
    __int128 v = (__int128)(u << 32) >> 32;
    v += carry;
    const unsigned __int128 dig = (unsigned __int128)v & mask;
    carry = v >> cb;
    // write dig (cb bits) at bit position pos
    size_t li = pos / 64;
    int sh = int(pos % 64);
    unsigned __int128 sh_d = dig << sh;
    limbs[li] |= uint64_t(sh_d);
    limbs[li + 1] |= uint64_t(sh_d >> 64);
    if (sh + cb > 128) limbs[li + 2] |= uint64_t(dig >> (128 - sh));
    pos += size_t(cb);
  }
  mpz_import(z, nl, -1, sizeof(uint64_t), 0, 0, limbs.data());
  if (carry != 0) {
    mpz_t t;
    mpz_init(t);
    const int64_t cc = int64_t(carry);
    mpz_set_si(t, cc);
    mpz_mul_2exp(t, t, nbits);
    mpz_add(z, z, t);
    mpz_clear(t);
  }
}

static uint32_t* slot_reg(Ctx& c, int r, size_t slot) { return c.reg(r) + slot * c.slot_stride; }

// Coefficients of x in base B = 2^cbits as DIGIT_LIMBS planes of 32-bit limbs
// (bits [i*cbits, (i+1)*cbits) of x, non-negative).
static void mpz_to_digit_planes(const Ctx& c, const mpz_t x, int cb, std::vector<uint32_t>& out) {
  const size_t n = c.n;
  out.assign(size_t(DIGIT_LIMBS) * n, 0);
  const size_t xs = mpz_size(x);
  const mp_limb_t* xp = mpz_limbs_read(x);
  auto bit32 = [&](size_t pos, int nb) -> uint32_t {  // nb <= 32 bits starting at pos
    const size_t li = pos / 64;
    const int sh = int(pos % 64);
    const uint64_t lo = li < xs ? uint64_t(xp[li]) : 0, hi = (li + 1) < xs ? uint64_t(xp[li + 1]) : 0;
    uint64_t r = lo >> sh;
    if (sh) r |= hi << (64 - sh);
    return uint32_t(nb == 32 ? r : (r & ((uint64_t(1) << nb) - 1)));
  };
  for (size_t i = 0; i < n; ++i) {
    const size_t pos = i * size_t(cb);
    for (int l = 0; l < DIGIT_LIMBS && 32 * l < cb; ++l) {
      const int nb = std::min(32, cb - 32 * l);
      out[size_t(l) * n + i] = bit32(pos + 32 * size_t(l), nb);
    }
  }
}

// set_mpz (src/transformGPU.h:323-360): the host splits x into digits, the
// device encodes them into K residue planes.
static int upload_mpz(Ctx& c, int r, const mpz_t x, size_t slot = 0) {
  std::vector<uint32_t> h;
  mpz_to_digit_planes(c, x, slot_cbits(c, slot), h);
  uint32_t* digits = c.rawtmp() + slot * c.slot_stride;
  HIPCHK(hipMemcpyAsync(digits, h.data(), h.size() * 4, hipMemcpyHostToDevice, c.stream));
  CarryArgs A = c.cbase;
  const dim3 grid(unsigned((c.n / CARRY_L + 255) / 256)), block(256);
  uint32_t* dst = slot_reg(c, r, slot);
  switch (c.pl.K) {
    case 2: hipLaunchKernelGGL(k_digits_to_res<2>, grid, block, 0, c.stream, digits, dst, A); break;
    case 3: hipLaunchKernelGGL(k_digits_to_res<3>, grid, block, 0, c.stream, digits, dst, A); break;
    case 4: hipLaunchKernelGGL(k_digits_to_res<4>, grid, block, 0, c.stream, digits, dst, A); break;
    default: hipLaunchKernelGGL(k_digits_to_res<5>, grid, block, 0, c.stream, digits, dst, A); break;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c.stream));
  return ANNIE_OK;
}

static int scale_reg(Ctx& c, int r, size_t slot = 0) {
  uint32_t s[KMAX] = {0, 0, 0, 0, 0};
  for (int k = 0; k < c.pl.K; ++k) s[k] = c.pc.sq_scale[k];  // mont(x, n^-1 2^64) = x n^-1 2^32
  const dim3 grid(unsigned((c.n + 255) / 256), c.pl.K), block(256);
  hipLaunchKernelGGL(k_scale, grid, block, 0, c.stream, slot_reg(c, r, slot), c.n, c.pl.K, c.nbase, s[0], s[1], s[2],
                     s[3], s[4]);
  HIPCHK(hipGetLastError());
  return ANNIE_OK;
}

// get() for every slot of the current list (slot 0 without a list): T_Q = src
// as raw digits (a carry pass over its own residues), T_R = src, reduce to
// raw digits, read back, digits -> integer mod N.  mods[i] is the modulus of
// list position i; zs[i] receives the value.
static int get_values(Ctx& c, int src, const size_t* slots, const mpz_t* const* mods, mpz_t* const* zs, size_t count) {
  int rc;
  const CarryJob cj = cjob(c.reg(src), c.tq(), 1u, 1, 0, 1);
  if ((rc = launch_carry(c, &cj, 1))) return rc;
  if ((rc = copy_reg(c, RTR, src))) return rc;
  if ((rc = reduce(c, c.rawtmp(), true))) return rc;
  const size_t w = size_t(DIGIT_LIMBS) * c.n;
  std::vector<uint32_t> d(w * count);
  for (size_t i = 0; i < count; ++i)
    HIPCHK(hipMemcpyAsync(d.data() + i * w, c.rawtmp() + slots[i] * c.slot_stride, w * 4, hipMemcpyDeviceToHost,
                          c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  for (size_t i = 0; i < count; ++i) {
    const std::vector<uint32_t> di(d.begin() + i * w, d.begin() + (i + 1) * w);
    digits_to_mpz(c, di, slot_cbits(c, slots[i]), *zs[i]);
    mpz_mod(*zs[i], *zs[i], *mods[i]);
  }
  return ANNIE_OK;
}

static int get_value(Ctx& c, int src, mpz_t z) {
  const size_t slot0 = 0;
  const mpz_t* m = &c.N;
  Mpz v;
  mpz_t* zp = &v.v;
  const int rc = get_values(c, src, &slot0, &m, &zp, 1);
  if (rc == ANNIE_OK) mpz_set(z, v.v);
  return rc;
}

static void limbs_to_mpz(mpz_t z, const uint64_t* l, size_t nl) { mpz_import(z, nl, -1, sizeof(uint64_t), 0, 0, l); }

static int mpz_to_limbs(const mpz_t z, uint64_t* out, size_t cap, size_t* nout) {
  const size_t need = (mpz_sizeinbase(z, 2) + 63) / 64;
  if (nout) *nout = mpz_sgn(z) == 0 ? 0 : need;
  if (mpz_sgn(z) == 0) return ANNIE_OK;
  if (need > cap) return fail(ANNIE_ERR_ARG, "output buffer too small");
  size_t cnt = 0;
  mpz_export(out, &cnt, -1, sizeof(uint64_t), 0, 0, z);
  for (size_t i = cnt; i < need; ++i) out[i] = 0;
  return ANNIE_OK;
}

static Plan plan_for(const mpz_t N, int32_t a) {
  long e = 0;
  const double m = mpz_get_d_2exp(&e, N);
  return select_plan(m, e, a);
}

static void destroy(annie_hip_t* h) {
  if (!h) return;
  Ctx& c = h->c;
  if (c.stream) {
    (void)hipSetDevice(c.device);
    (void)hipStreamSynchronize(c.stream);
  }
  if (c.aux) {
    (void)hipStreamSynchronize(c.aux);
    (void)hipStreamDestroy(c.aux);
  }
  if (c.ev_fork) (void)hipEventDestroy(c.ev_fork);
  if (c.ev_join) (void)hipEventDestroy(c.ev_join);
  if (c.mem) (void)hipFree(c.mem);
  if (c.tw) (void)hipFree(c.tw);
  for (auto& g : c.graphs) (void)hipGraphExecDestroy(g.exec);
  if (c.graph_buf) (void)hipFree(c.graph_buf);
  if (c.cbits_dev) (void)hipFree(c.cbits_dev);
  for (Stage* st : {&c.stage, &c.run_stage}) {
    if (st->dev) (void)hipFree(st->dev);
    if (st->host) (void)hipHostFree(st->host);
    for (hipEvent_t e : st->ev)
      if (e) (void)hipEventDestroy(e);
  }
  for (hipEvent_t e : c.ev) (void)hipEventDestroy(e);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  mpz_clears(c.N, c.factor, nullptr);
  delete h;
}

// R = B^n - 1 for B = 2^cbits, n = 2^LN.
static void plan_R(int cbits, int LN, mpz_t R) {
  mpz_set_ui(R, 1);
  mpz_mul_2exp(R, R, mp_bitcnt_t(cbits) << LN);
  mpz_sub_ui(R, R, 1);
}

static void plan_R(const Plan& pl, mpz_t R) { plan_R(pl.cbits(), pl.LN, R); }

// Device state shared by the slots of a context: stream, register file of
// `slots` slots, twiddle tables, kernel constants.
static int setup_device(Ctx& c, int device, size_t slots) {
  c.n = c.pl.n();
  c.m1 = (c.pl.LN <= 13) ? 0 : c.pl.LN / 2;
  c.m2 = c.pl.LN - c.m1;
  c.device = device;
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  c.cur = c.stream;
  // ANNIE_SINGLE_STREAM=1 keeps both halves of a squaring on one stream.
  // The two-stream squaring costs the host ~0.08-0.11 ms per squaring to
  // enqueue (16 launches and 4 event calls, against ~0.04 ms for 12 launches
  // on one stream) and gains only what the second stream fills of the GPU, so
  // it is used from K n >= 2^22 words per register.  Measured ms/squaring,
  // one vs two streams: K = 3 at 2^16 0.071 / 0.086, 2^17 0.079 / 0.092, 2^18
  // 0.100 / 0.109, 2^19 0.133 / 0.134; K = 5 at 2^21 0.515 / 0.495.
  const char* single = std::getenv("ANNIE_SINGLE_STREAM");
  if (single && single[0] == '1') c.dual_stream = false;
  if ((size_t(c.pl.K) << c.pl.LN) < (size_t(1) << 22)) c.dual_stream = false;
  if (slots == 1 && c.m1 > 0 && c.dual_stream) {
    HIPCHK(hipStreamCreateWithFlags(&c.aux, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
  }
  const size_t K = size_t(c.pl.K), n = c.n;
  c.nslots_alloc = slots;
  c.slot_stride = size_t(NREG) * K * n + 2 * size_t(DIGIT_LIMBS) * n;
  if (hipMalloc(&c.mem, slots * c.slot_stride * 4) != hipSuccess) return fail(ANNIE_ERR_NOMEM, "register file");
  c.regs = c.mem;
  c.raw = c.mem + size_t(NREG) * K * n;
  if (hipMalloc(&c.tw, 4 * K * n * sizeof(TwDev)) != hipSuccess) return fail(ANNIE_ERR_NOMEM, "twiddles");
  HIPCHK(hipMemsetAsync(c.mem, 0, slots * c.slot_stride * 4, c.stream));

  c.pc = make_consts(c.pl);
  for (int k = 0; k < KMAX; ++k) {
    c.nbase.p[k] = c.pc.p[k];
    c.nbase.q[k] = c.pc.q[k];
    c.nbase.sc[k] = c.pc.sq_scale[k];
    c.nbase.scq[k] = c.pc.sq_scale_q[k];
  }
  c.nbase.LN = c.cbase.LN = c.pl.LN;
  c.cbase.cbits = c.pl.cbits();
  c.nbase.njobs = c.cbase.njobs = 1;

  std::vector<TwDev> host(4 * K * n);
  for (size_t k = 0; k < K; ++k) {
    const PrimeTables pt = make_tables(c.pl.LN, int(k));
    for (size_t i = 0; i < n; ++i) {
      host[(0 * K + k) * n + i] = pt.fwdR[i].w;
      host[(1 * K + k) * n + i] = pt.invR[i].w;
      host[(2 * K + k) * n + i] = pt.fwdQ[i].w;
      host[(3 * K + k) * n + i] = pt.invQ[i].w;
    }
  }
  HIPCHK(hipMemcpyAsync(c.tw, host.data(), host.size() * sizeof(TwDev), hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return ANNIE_OK;
}

// Per-modulus constants of one slot: N_Q, the negacyclic transform of N, and
// Ninv_R, the cyclic transform of N^-1 mod R (src/transformGPU.h:737-741),
// both scaled by n^-1 2^32 for the pointwise step, and the MMGW form of 1,
// R/2 mod N (src/transformGPU.h:743-747).  Ninv is N^-1 mod R.
static int load_slot_consts(Ctx& c, size_t slot, const mpz_t N, const mpz_t Ninv) {
  int rc;
  const uint32_t* prev_ids = c.slot_ids;
  const int prev_n = c.nslots;
  c.slot_ids = nullptr;  // the transforms below address the slot directly
  c.nslots = 1;
  const size_t off = slot * c.slot_stride;
  auto done = [&](int r) {
    c.slot_ids = prev_ids;
    c.nslots = prev_n;
    return r;
  };
  if ((rc = upload_mpz(c, RTMP, N, slot))) return done(rc);
  {
    NttJob j = job(c.reg(RTMP) + off, c.reg(RNQ) + off, c.fwdQ(), c.invQ());
    if ((rc = forward(c, &j, 1))) return done(rc);
    if ((rc = scale_reg(c, RNQ, slot))) return done(rc);
  }
  if ((rc = upload_mpz(c, RTMP, Ninv, slot))) return done(rc);
  {
    NttJob j = job(c.reg(RTMP) + off, c.reg(RNIR) + off, c.fwdR(), c.invR());
    if ((rc = forward(c, &j, 1))) return done(rc);
    if ((rc = scale_reg(c, RNIR, slot))) return done(rc);
  }
  mpz_t R, t;
  mpz_inits(R, t, nullptr);
  plan_R(slot_cbits(c, slot), c.pl.LN, R);
  mpz_mod(t, R, N);
  if (mpz_odd_p(t)) mpz_add(t, t, N);
  mpz_fdiv_q_2exp(t, t, 1);
  rc = upload_mpz(c, RONE, t, slot);
  mpz_clears(R, t, nullptr);
  if (rc) return done(rc);
  HIPCHK(hipStreamSynchronize(c.stream));
  return done(ANNIE_OK);
}

static int create(const uint64_t* Nl, size_t nl, int32_t a, int device, int min_ln, annie_hip_t** out) {
  if (!out || !Nl || nl == 0) return fail(ANNIE_ERR_ARG, "null argument");
  *out = nullptr;
  if (a < 1) return fail(ANNIE_ERR_ARG, "base must be positive");
  std::unique_ptr<annie_hip, void (*)(annie_hip_t*)> h(new annie_hip(), destroy);
  Ctx& c = h->c;
  mpz_inits(c.N, c.factor, nullptr);
  limbs_to_mpz(c.N, Nl, nl);
  if (mpz_even_p(c.N) || mpz_cmp_ui(c.N, 3) < 0) return fail(ANNIE_ERR_ARG, "N must be odd and >= 3");
  c.pl = plan_for(c.N, a);
  if (c.pl.LN < min_ln) return fail(ANNIE_ERR_SIZE, "NTT size < 2^" + std::to_string(min_ln) + " is not supported");
  if (c.pl.LN > 23) return fail(ANNIE_ERR_SIZE, "NTT size > 2^23 is not supported");
  if (c.pl.LN < 4) return fail(ANNIE_ERR_SIZE, "NTT size < 2^4 is not supported by this engine");

  // R = B^n - 1; N must be invertible mod R (src/transformGPU.h:652-661).
  mpz_t R, t;
  mpz_inits(R, t, nullptr);
  struct Clear {
    mpz_t& a;
    mpz_t& b;
    ~Clear() { mpz_clears(a, b, nullptr); }
  } clear_rt{R, t};
  plan_R(c.pl, R);
  if (mpz_invert(t, c.N, R) == 0) {
    mpz_gcd(c.factor, c.N, R);
    c.init_failed = true;
    *out = h.release();
    return ANNIE_INIT_FAILED;
  }
  int rc;
  if ((rc = setup_device(c, device, 1))) return rc;
  if ((rc = load_slot_consts(c, 0, c.N, t))) return rc;
  *out = h.release();
  return ANNIE_OK;
}

// Uploads a slot list (and optionally per-slot values) through the staging
// ring; returns device pointers valid for operations enqueued afterwards.
static int stage_upload(Ctx& c, const uint32_t* ids, const uint32_t* vals, size_t count, const uint32_t** d_ids,
                        const uint32_t** d_vals, uint32_t* dst = nullptr) {
  Stage& st = c.stage;
  const size_t words = count * (vals ? 2 : 1);
  if (words > st.cap) return fail(ANNIE_ERR_ARG, "slot list longer than the batch");
  const int e = st.pos;
  st.pos = (st.pos + 1) % NSTAGE;
  if (st.pending[e]) {
    HIPCHK(hipEventSynchronize(st.ev[e]));
    st.pending[e] = false;
  }
  uint32_t* h = st.host + size_t(e) * st.cap;
  uint32_t* d = dst ? dst : st.dev + size_t(e) * st.cap;
  std::memcpy(h, ids, count * 4);
  if (vals) std::memcpy(h + count, vals, count * 4);
  HIPCHK(hipMemcpyAsync(d, h, words * 4, hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipEventRecord(st.ev[e], c.stream));
  st.pending[e] = true;
  *d_ids = d;
  if (d_vals) *d_vals = vals ? d + count : nullptr;
  return ANNIE_OK;
}

// The slot list of one batch operation.
struct SlotScope {
  Ctx& c;
  SlotScope(Ctx& cc, const uint32_t* ids, size_t count) : c(cc) {
    c.slot_ids = ids;
    c.nslots = int(count);
  }
  ~SlotScope() {
    c.slot_ids = nullptr;
    c.nslots = 1;
  }
};

// Slot lists: non-empty, ids in range, each id once.
static int check_slots(const Ctx& c, const uint32_t* ids, size_t count) {
  if (!ids || count == 0 || count > c.nslots_alloc) return fail(ANNIE_ERR_ARG, "bad slot list");
  std::vector<uint8_t> seen(c.nslots_alloc, 0);
  for (size_t i = 0; i < count; ++i) {
    if (ids[i] >= c.nslots_alloc) return fail(ANNIE_ERR_ARG, "slot id out of range");
    if (seen[ids[i]]++) return fail(ANNIE_ERR_ARG, "slot listed twice");
    if (!c.slotN[ids[i]]) return fail(ANNIE_ERR_ARG, "slot has no modulus loaded");
  }
  return ANNIE_OK;
}

static bool ready(const annie_hip_t* h) {
  if (!h) {
    g_last_error = "null context";
    return false;
  }
  if (h->c.init_failed) {
    g_last_error = "context failed to initialise (gcd(N, R) > 1)";
    return false;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Squaring without allocation: the work registers of one squaring (four
// registers of K n words and the T_Q digits) in a file-scope device array,
// one copy per GPU, shared by every call on that GPU.  In host code
// `scratch` only names the symbol; its device address on each GPU comes from
// hipGetSymbolAddress, looked up once per device with that device current.
constexpr size_t NOALLOC_KN_MAX = ANNIE_HIP_NOALLOC_MAX_KN;
constexpr size_t SCRATCH_WORDS = 4 * NOALLOC_KN_MAX + size_t(DIGIT_LIMBS) * (NOALLOC_KN_MAX / 2);
static __device__ __align__(256) uint32_t scratch[SCRATCH_WORDS];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static uint32_t* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device `dev`.
static int scratch_on(int dev, uint32_t** out) {
  if (dev < 0 || dev >= MAX_DEVICES) return fail(ANNIE_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return fail(ANNIE_ERR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return ANNIE_OK;
}

}  // namespace annie

using namespace annie;

extern "C" {

int annie_hip_select_plan(const uint64_t* Nl, size_t nl, int32_t a, int* K, int* pck, int* LN, int* lbase) {
  if (!Nl || nl == 0 || a < 1) return fail(ANNIE_ERR_ARG, "bad argument");
  mpz_t N;
  mpz_init(N);
  limbs_to_mpz(N, Nl, nl);
  const Plan pl = plan_for(N, a);
  mpz_clear(N);
  if (K) *K = pl.K;
  if (pck) *pck = pl.pck;
  if (LN) *LN = pl.LN;
  if (lbase) *lbase = pl.lbase;
  return ANNIE_OK;
}

int annie_hip_plan_constants(int LN, int K, int pck, int lbase, uint32_t* out, size_t cap) {
  if (!out || cap < 5 * KMAX + KMAX * KMAX || LN < 1 || LN > 24) return fail(ANNIE_ERR_ARG, "bad argument");
  Plan pl;
  pl.K = K;
  pl.pck = pck;
  pl.LN = LN;
  pl.lbase = lbase;
  const PrimeConsts c = make_consts(pl);
  for (int k = 0; k < KMAX; ++k) {
    out[k] = c.p[k];
    out[KMAX + k] = c.q[k];
    out[2 * KMAX + k] = c.sq_scale[k];
    out[3 * KMAX + k] = c.c32[k];
    out[4 * KMAX + k] = c.c64[k];
    for (int j = 0; j < KMAX; ++j) out[5 * KMAX + k * KMAX + j] = (j == k) ? 0u : c.ginv[k][j];
  }
  return ANNIE_OK;
}

int annie_hip_twiddles(int LN, int k, int table, uint32_t* w, uint32_t* wq, size_t cap) {
  if (LN < 1 || LN > 23 || k < 0 || k >= KMAX || table < 0 || table > 3 || !w || !wq || cap < (size_t(1) << LN))
    return fail(ANNIE_ERR_ARG, "bad argument");
  const PrimeTables t = make_tables(LN, k);
  const std::vector<Tw>& v = table == 0 ? t.fwdR : table == 1 ? t.invR : table == 2 ? t.fwdQ : t.invQ;
  for (size_t i = 0; i < v.size(); ++i) {
    w[i] = v[i].w;
    wq[i] = v[i].wq;
  }
  return ANNIE_OK;
}

int annie_hip_create(const uint64_t* N, size_t nlimbs, int32_t a, int device, int min_ln, annie_hip_t** out) {
  annie_hip_t* h = nullptr;
  const int rc = annie::create(N, nlimbs, a, device, min_ln, &h);
  if (rc < 0) {
    // create() leaves no context behind on error.
    if (out) *out = nullptr;
    return rc;
  }
  *out = h;
  return rc;
}

int annie_hip_get_factor(const annie_hip_t* h, uint64_t* limbs, size_t cap, size_t* nout) {
  if (!h) return fail(ANNIE_ERR_ARG, "null context");
  const mpz_t& f = h->c.init_failed ? h->c.factor : h->c.N;
  return mpz_to_limbs(f, limbs, cap, nout);
}

int annie_hip_set1(annie_hip_t* h) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  return copy_reg(h->c, R0, RONE);
}

int annie_hip_copy(annie_hip_t* h, size_t dst, size_t src) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  if (dst > 2 || src > 2) return fail(ANNIE_ERR_ARG, "register index out of range");
  return copy_reg(h->c, int(dst), int(src));
}

int annie_hip_square_mul(annie_hip_t* h, int32_t a) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  if (a < 1) return fail(ANNIE_ERR_ARG, "multiplier must be positive");
  return annie::square_mul(h->c, uint32_t(a));
}

int annie_hip_square_mul_noalloc(annie_hip_t* h, const uint32_t* x, uint32_t* y, int32_t a) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  Ctx& c = h->c;
  if (a < 1) return fail(ANNIE_ERR_ARG, "multiplier must be positive");
  if (!x || !y) return fail(ANNIE_ERR_ARG, "null register buffer");
  if ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 15u)
    return fail(ANNIE_ERR_ARG, "register buffers must be 16-byte aligned");
  if (c.nslots_alloc != 1) return fail(ANNIE_ERR_ARG, "not a single-modulus context");
  const size_t kn = size_t(c.pl.K) * c.n;
  if (kn > NOALLOC_KN_MAX || 2 * c.n > NOALLOC_KN_MAX)
    return fail(ANNIE_ERR_SIZE, "transform of " + std::to_string(kn) + " words per register exceeds the no-allocation "
                                "limit of " + std::to_string(NOALLOC_KN_MAX));
  uint32_t* w = nullptr;
  HIPCHK(hipSetDevice(c.device));
  int rc;
  if ((rc = scratch_on(c.device, &w))) return rc;
  const SqBufs B{x, y, w, w + kn, w + 2 * kn, w + 3 * kn, w + 4 * kn};
  // One stream: the two-stream squaring is only used above this size limit.
  hipStream_t aux = c.aux;
  c.aux = nullptr;
  rc = annie::square_mul(c, B, uint32_t(a));
  c.aux = aux;
  return rc;
}

int annie_hip_square_mul_bits(annie_hip_t* h, const uint8_t* bits, size_t count, int32_t a) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  if (a < 1 || (count && !bits)) return fail(ANNIE_ERR_ARG, "bad argument");
  for (size_t i = 0; i < count; ++i) {
    const int rc = annie::square_mul(h->c, bits[i] ? uint32_t(a) : 1u);
    if (rc) return rc;
  }
  return ANNIE_OK;
}

int annie_hip_mul(annie_hip_t* h, size_t dst, size_t src) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  if (dst > 2 || src > 2) return fail(ANNIE_ERR_ARG, "register index out of range");
  return annie::mul(h->c, int(dst), int(src));
}

int annie_hip_get(annie_hip_t* h, size_t src, uint64_t* limbs, size_t cap, size_t* nout) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  if (src > 2) return fail(ANNIE_ERR_ARG, "register index out of range");
  mpz_t z;
  mpz_init(z);
  int rc = get_value(h->c, int(src), z);
  if (rc == ANNIE_OK) rc = mpz_to_limbs(z, limbs, cap, nout);
  mpz_clear(z);
  return rc;
}

size_t annie_hip_checkpoint_size(const annie_hip_t* h) {
  if (!h) return 0;
  return 8 + 2 * size_t(h->c.pl.K) * h->c.n * 4;
}

int annie_hip_save_checkpoint(annie_hip_t* h, void* buf, size_t cap) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  Ctx& c = h->c;
  if (!buf || cap < annie_hip_checkpoint_size(h)) return fail(ANNIE_ERR_ARG, "checkpoint buffer too small");
  uint32_t* w = static_cast<uint32_t*>(buf);
  w[0] = uint32_t(c.pl.K);
  w[1] = uint32_t(c.n);
  for (int k = 0; k < c.pl.K; ++k) {
    HIPCHK(hipMemcpyAsync(w + 2 + size_t(2 * k) * c.n, c.reg(R0) + size_t(k) * c.n, c.n * 4, hipMemcpyDeviceToHost,
                          c.stream));
    HIPCHK(hipMemcpyAsync(w + 2 + size_t(2 * k + 1) * c.n, c.reg(R1) + size_t(k) * c.n, c.n * 4,
                          hipMemcpyDeviceToHost, c.stream));
  }
  HIPCHK(hipStreamSynchronize(c.stream));
  return ANNIE_OK;
}

int annie_hip_load_checkpoint(annie_hip_t* h, const void* buf, size_t size) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  Ctx& c = h->c;
  if (!buf || size < annie_hip_checkpoint_size(h)) return fail(ANNIE_ERR_ARG, "checkpoint too short");
  const uint32_t* w = static_cast<const uint32_t*>(buf);
  if (w[0] != uint32_t(c.pl.K) || w[1] != uint32_t(c.n)) return fail(ANNIE_ERR_ARG, "checkpoint of another transform");
  for (int k = 0; k < c.pl.K; ++k) {
    HIPCHK(hipMemcpyAsync(c.reg(R0) + size_t(k) * c.n, w + 2 + size_t(2 * k) * c.n, c.n * 4, hipMemcpyHostToDevice,
                          c.stream));
    HIPCHK(hipMemcpyAsync(c.reg(R1) + size_t(k) * c.n, w + 2 + size_t(2 * k + 1) * c.n, c.n * 4,
                          hipMemcpyHostToDevice, c.stream));
  }
  HIPCHK(hipStreamSynchronize(c.stream));
  return ANNIE_OK;
}

int annie_hip_info(const annie_hip_t* h, int* K, int* pck, int* LN, int* lbase) {
  if (!h) return fail(ANNIE_ERR_ARG, "null context");
  if (K) *K = h->c.pl.K;
  if (pck) *pck = h->c.pl.pck;
  if (LN) *LN = h->c.pl.LN;
  if (lbase) *lbase = h->c.pl.lbase;
  return ANNIE_OK;
}

static void plan_levels(int LN, int* m1, int* m2) {
  *m1 = (LN <= 13) ? 0 : LN / 2;
  *m2 = LN - *m1;
}

int annie_hip_schedule_fits(int LN, int row_lines, int col_cols) {
  if (LN < 4 || LN > 23 || row_lines < 0 || col_cols < 0) return 0;
  int m1, m2;
  plan_levels(LN, &m1, &m2);
  return schedule_fits(m1, m2, LN, uint32_t(row_lines), uint32_t(col_cols)) ? 1 : 0;
}

int annie_hip_set_schedule(annie_hip_t* h, int row_lines, int col_cols) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  Ctx& c = h->c;
  if (row_lines < 0 || col_cols < 0 || !schedule_fits(c.m1, c.m2, c.pl.LN, uint32_t(row_lines), uint32_t(col_cols)))
    return fail(ANNIE_ERR_ARG, "schedule does not fit this transform length");
  HIPCHK(hipStreamSynchronize(c.stream));
  if (c.aux) HIPCHK(hipStreamSynchronize(c.aux));
  for (auto& g : c.graphs) (void)hipGraphExecDestroy(g.exec);
  c.graphs.clear();
  c.sched_lines = uint32_t(row_lines);
  c.sched_cols = uint32_t(col_cols);
  return ANNIE_OK;
}

int annie_hip_get_schedule(const annie_hip_t* h, int* row_lines, int* col_cols) {
  if (!h) return fail(ANNIE_ERR_ARG, "null context");
  const Ctx& c = h->c;
  const bool short_rows = (1u << c.m2) / NTT_R < 64;
  const bool narrow_cols = c.m1 > 0 && c.m1 < 10;
  if (row_lines) *row_lines = short_rows ? int(c.sched_lines ? c.sched_lines : row_lines_default(c.m1, c.m2)) : 0;
  if (col_cols) *col_cols = narrow_cols ? int(c.sched_cols ? c.sched_cols : col_cols_default(c.m1, c.pl.LN)) : 0;
  return ANNIE_OK;
}

int annie_hip_sync(annie_hip_t* h) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  HIPCHK(hipStreamSynchronize(h->c.stream));
  return ANNIE_OK;
}

int annie_hip_read_register(annie_hip_t* h, size_t reg, uint32_t* out, size_t cap_words) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  Ctx& c = h->c;
  const size_t words = size_t(c.pl.K) * c.n;
  if (reg >= size_t(NREG) || !out || cap_words < words) return fail(ANNIE_ERR_ARG, "bad register read");
  HIPCHK(hipMemcpyAsync(out, c.reg(int(reg)), words * 4, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return ANNIE_OK;
}

int annie_hip_write_register(annie_hip_t* h, size_t reg, const uint32_t* in, size_t words) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  Ctx& c = h->c;
  if (reg > 2 || !in || words != size_t(c.pl.K) * c.n) return fail(ANNIE_ERR_ARG, "bad register write");
  HIPCHK(hipMemcpyAsync(c.reg(int(reg)), in, words * 4, hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return ANNIE_OK;
}

// Squarings of annie_hip_square_mul_bits timed with HIP events on the
// context's stream (the elapsed GPU time of the whole sequence, in ms).
int annie_hip_time_square_mul_bits(annie_hip_t* h, const uint8_t* bits, size_t count, int32_t a, float* ms) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  if (a < 1 || (count && !bits) || !ms) return fail(ANNIE_ERR_ARG, "bad argument");
  Ctx& c = h->c;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c.stream));
  int rc = ANNIE_OK;
  for (size_t i = 0; i < count && rc == ANNIE_OK; ++i) rc = annie::square_mul(c, bits[i] ? uint32_t(a) : 1u);
  HIPCHK(hipEventRecord(e1, c.stream));
  HIPCHK(hipEventSynchronize(e1));
  HIPCHK(hipEventElapsedTime(ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// Per-kernel timing: begin clears and enables event pairs around every launch;
// end synchronises and writes one line per kernel kind, "name count total_ms".
int annie_hip_profile_begin(annie_hip_t* h) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  h->c.profiling = true;
  h->c.rec_kind.clear();
  return ANNIE_OK;
}

int annie_hip_profile_end(annie_hip_t* h, char* buf, size_t cap) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  Ctx& c = h->c;
  c.profiling = false;
  HIPCHK(hipStreamSynchronize(c.stream));
  std::vector<double> tot(c.kinds.size(), 0.0);
  std::vector<long> cnt(c.kinds.size(), 0);
  for (size_t r = 0; r < c.rec_kind.size(); ++r) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c.ev[2 * r], c.ev[2 * r + 1]));
    tot[c.rec_kind[r]] += ms;
    cnt[c.rec_kind[r]] += 1;
  }
  std::string out;
  for (size_t k = 0; k < c.kinds.size(); ++k) {
    if (!cnt[k]) continue;
    char line[256];
    snprintf(line, sizeof(line), "%s %ld %.6f\n", c.kinds[k].c_str(), cnt[k], tot[k]);
    out += line;
  }
  c.rec_kind.clear();
  if (!buf || cap == 0) return fail(ANNIE_ERR_ARG, "null buffer");
  std::strncpy(buf, out.c_str(), cap - 1);
  buf[cap - 1] = 0;
  return ANNIE_OK;
}

// Diagnostics: run the squaring r0 -> a r0^2 up to `stage` and leave the
// intermediate in place (1: convolution residues in registers 5 (cyclic) and
// 6 (negacyclic); 2: T_R residues in register 4 and T_Q digits; 3: the m'
// convolution in register 5; 4: m' + 2 in register 4).  r0 is not updated.
int annie_hip_debug_square_stage(annie_hip_t* h, int32_t a, int stage) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  Ctx& c = h->c;
  int rc;
  const NttJob j[2] = {job(c.reg(R0), c.reg(RWR), c.fwdR(), c.invR()), job(c.reg(R0), c.reg(RWQ), c.fwdQ(), c.invQ())};
  if ((rc = convolve(c, ROW_SQR, j, 2))) return rc;
  if (stage >= 2) {
    const CarryJob cj[2] = {cjob(c.reg(RWR), c.reg(RTR), uint32_t(a), 0, 0, 0), cjob(c.reg(RWQ), c.tq(), uint32_t(a), 1, 0, 1)};
    if ((rc = launch_carry(c, cj, 2))) return rc;
  }
  if (stage >= 3) {
    NttJob jr = job(c.reg(RTR), c.reg(RWR), c.fwdR(), c.invR(), c.reg(RNIR));
    if ((rc = convolve(c, ROW_MULC, &jr, 1))) return rc;
  }
  if (stage >= 4) {
    CarryJob cj = cjob(c.reg(RWR), c.reg(RTR), 1u, 0, 1, 0);
    if ((rc = launch_carry(c, &cj, 1))) return rc;
  }
  HIPCHK(hipStreamSynchronize(c.stream));
  return ANNIE_OK;
}

// Diagnostics: the T_Q digit planes (DIGIT_LIMBS planes of n words).
int annie_hip_read_tq(annie_hip_t* h, uint32_t* out, size_t cap_words) {
  if (!ready(h)) return ANNIE_ERR_ARG;
  Ctx& c = h->c;
  if (!out || cap_words < size_t(DIGIT_LIMBS) * c.n) return fail(ANNIE_ERR_ARG, "buffer too small");
  HIPCHK(hipMemcpyAsync(out, c.tq(), size_t(DIGIT_LIMBS) * c.n * 4, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return ANNIE_OK;
}

// ---------------------------------------------------------------------------
// Batches: many moduli of one length class (one plan) in one allocation; every
// operation runs for a list of slots with one set of launches.

int annie_hip_batch_create(int K, int pck, int LN, int lbase, size_t slots, int device, annie_hip_t** out) {
  if (!out) return fail(ANNIE_ERR_ARG, "null argument");
  *out = nullptr;
  if (K < 2 || K > KMAX || (pck != 1 && pck != 2 && pck != 4) || LN < 4 || LN > 23 || lbase < 2 || lbase > 31 ||
      pck * lbase > 32 * DIGIT_LIMBS - 1)
    return fail(ANNIE_ERR_ARG, "bad plan");
  if (slots == 0 || slots > 32767) return fail(ANNIE_ERR_ARG, "slot count must be in [1, 32767]");
  std::unique_ptr<annie_hip, void (*)(annie_hip_t*)> h(new annie_hip(), destroy);
  Ctx& c = h->c;
  mpz_inits(c.N, c.factor, nullptr);
  c.pl.K = K;
  c.pl.pck = pck;
  c.pl.LN = LN;
  c.pl.lbase = lbase;
  int rc;
  if ((rc = setup_device(c, device, slots))) return rc;
  c.slotN.resize(slots);
  c.slot_cb.assign(slots, c.pl.cbits());
  c.slot_base.assign(slots, 1u);
  if (hipMalloc(&c.cbits_dev, 3 * slots * 4) != hipSuccess) return fail(ANNIE_ERR_NOMEM, "per-slot digit widths");
  c.base_dev = c.cbits_dev + slots;
  c.list_dev = c.cbits_dev + 2 * slots;
  {
    std::vector<uint32_t> cb(3 * slots, 1u);
    std::fill(cb.begin(), cb.begin() + slots, uint32_t(c.pl.cbits()));
    HIPCHK(hipMemcpy(c.cbits_dev, cb.data(), 3 * slots * 4, hipMemcpyHostToDevice));
  }
  c.cbase.cbits_slot = c.cbits_dev;
  Stage& st = c.stage;
  st.cap = 2 * slots;
  if (hipMalloc(&st.dev, NSTAGE * st.cap * 4) != hipSuccess) return fail(ANNIE_ERR_NOMEM, "staging");
  if (hipHostMalloc(&st.host, NSTAGE * st.cap * 4) != hipSuccess) return fail(ANNIE_ERR_NOMEM, "pinned staging");
  for (int e = 0; e < NSTAGE; ++e) HIPCHK(hipEventCreateWithFlags(&st.ev[e], hipEventDisableTiming));
  *out = h.release();
  return ANNIE_OK;
}

int annie_hip_batch_bytes(int K, int LN, size_t* per_slot, size_t* fixed) {
  if (K < 2 || K > KMAX || LN < 4 || LN > 23) return fail(ANNIE_ERR_ARG, "bad plan");
  const size_t n = size_t(1) << LN;
  // setup_device's slot stride, the staging ring's two words per slot per
  // entry, the graph buffers' two words per slot and the per-slot width,
  // base and list words.
  if (per_slot)
    *per_slot = (size_t(NREG) * size_t(K) * n + 2 * size_t(DIGIT_LIMBS) * n) * 4 + NSTAGE * 2 * 4 + 2 * 4 + 3 * 4;
  if (fixed) *fixed = 4 * size_t(K) * n * sizeof(TwDev);
  return ANNIE_OK;
}

int annie_hip_batch_load(annie_hip_t* h, size_t slot, const uint64_t* N, size_t nlimbs, int32_t a, uint64_t* factor,
                         size_t cap, size_t* nout) {
  if (!h || !N || nlimbs == 0 || a < 1) return fail(ANNIE_ERR_ARG, "bad argument");
  Ctx& c = h->c;
  if (c.slotN.empty() || slot >= c.nslots_alloc) return fail(ANNIE_ERR_ARG, "not a batch, or slot out of range");
  auto m = std::make_unique<Mpz>();
  limbs_to_mpz(m->v, N, nlimbs);
  if (mpz_even_p(m->v) || mpz_cmp_ui(m->v, 3) < 0) return fail(ANNIE_ERR_ARG, "N must be odd and >= 3");
  const Plan pl = plan_for(m->v, a);
  // Same K, pck and LN; lbase may differ within the batch's digit-width
  // class (pck lbase >> 5, the carry kernel's limb shift).
  if (pl.K != c.pl.K || pl.pck != c.pl.pck || pl.LN != c.pl.LN || (pl.cbits() >> 5) != (c.pl.cbits() >> 5))
    return fail(ANNIE_ERR_SIZE, "N belongs to another length class");
  mpz_t R, t;
  mpz_inits(R, t, nullptr);
  plan_R(pl, R);
  if (mpz_invert(t, m->v, R) == 0) {
    mpz_gcd(t, m->v, R);
    const int r = mpz_to_limbs(t, factor, cap, nout);
    mpz_clears(R, t, nullptr);
    c.slotN[slot].reset();
    return r == ANNIE_OK ? ANNIE_INIT_FAILED : r;
  }
  HIPCHK(hipSetDevice(c.device));
  c.slot_cb[slot] = pl.cbits();
  c.slot_base[slot] = uint32_t(a);
  const uint32_t cb = uint32_t(pl.cbits()), ab = uint32_t(a);
  HIPCHK(hipMemcpyAsync(c.cbits_dev + slot, &cb, 4, hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipMemcpyAsync(c.base_dev + slot, &ab, 4, hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  const int rc = load_slot_consts(c, slot, m->v, t);
  mpz_clears(R, t, nullptr);
  if (rc) return rc;
  c.slotN[slot] = std::move(m);
  if (nout) *nout = 0;
  return ANNIE_OK;
}

int annie_hip_batch_set1(annie_hip_t* h, const uint32_t* slots, size_t count) {
  if (!h) return fail(ANNIE_ERR_ARG, "null context");
  Ctx& c = h->c;
  int rc;
  if ((rc = check_slots(c, slots, count))) return rc;
  const uint32_t* d;
  if ((rc = stage_upload(c, slots, nullptr, count, &d, nullptr))) return rc;
  SlotScope sc(c, d, count);
  return copy_reg(c, R0, RONE);
}

int annie_hip_batch_copy(annie_hip_t* h, size_t dst, size_t src, const uint32_t* slots, size_t count) {
  if (!h) return fail(ANNIE_ERR_ARG, "null context");
  Ctx& c = h->c;
  if (dst > 2 || src > 2) return fail(ANNIE_ERR_ARG, "register index out of range");
  int rc;
  if ((rc = check_slots(c, slots, count))) return rc;
  const uint32_t* d;
  if ((rc = stage_upload(c, slots, nullptr, count, &d, nullptr))) return rc;
  SlotScope sc(c, d, count);
  return copy_reg(c, int(dst), int(src));
}

// Executable graph of one batched squaring over `count` slots whose ids and
// multipliers are read from graph_buf (captured on first use).
static int square_graph(Ctx& c, size_t count, hipGraphExec_t* out) {
  for (auto& g : c.graphs)
    if (g.count == count) {
      g.used = ++c.graph_clock;
      *out = g.exec;
      return ANNIE_OK;
    }
  constexpr size_t MAX_GRAPHS = 8;
  if (c.graphs.size() >= MAX_GRAPHS) {
    auto lru = std::min_element(c.graphs.begin(), c.graphs.end(),
                                [](const Ctx::Graph& x, const Ctx::Graph& y) { return x.used < y.used; });
    HIPCHK(hipStreamSynchronize(c.stream));
    (void)hipGraphExecDestroy(lru->exec);
    c.graphs.erase(lru);
  }
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  int rc;
  {
    SlotScope sc(c, c.graph_buf, count);
    rc = annie::square_mul(c, 1u, c.graph_buf + count);
  }
  const hipError_t e = hipStreamEndCapture(c.stream, &graph);
  if (rc != ANNIE_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(ANNIE_ERR, std::string("graph capture: ") + hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return fail(ANNIE_ERR, std::string("graph instantiate: ") + hipGetErrorString(ei));
  c.graphs.push_back({count, exec, ++c.graph_clock});
  *out = exec;
  return ANNIE_OK;
}

int annie_hip_batch_square_mul(annie_hip_t* h, const uint32_t* slots, const uint32_t* a, size_t count) {
  if (!h || !a) return fail(ANNIE_ERR_ARG, "null argument");
  Ctx& c = h->c;
  int rc;
  if ((rc = check_slots(c, slots, count))) return rc;
  for (size_t i = 0; i < count; ++i)
    if ((a[i] & ~SLOT_SNAP) < 1) return fail(ANNIE_ERR_ARG, "multiplier must be in [1, 2^31)");
  const uint32_t *d, *da;
  if (!c.use_graphs) {
    // Short lists whose multipliers are 1 or the slot's base: the multipliers
    // and snapshot flags travel as bits in the carries' kernel arguments and
    // the slot list stays on the device until it changes, so a squaring in
    // the steady state issues no host-to-device copy.
    bool short_list = count <= size_t(INLINE_SLOTS);
    for (size_t i = 0; i < count && short_list; ++i) {
      const uint32_t m = a[i] & ~SLOT_SNAP;
      short_list = m == 1u || m == c.slot_base[slots[i]];
    }
    if (short_list) {
      if (c.list_host.size() != count || !std::equal(slots, slots + count, c.list_host.begin())) {
        if ((rc = stage_upload(c, slots, nullptr, count, &d, nullptr, c.list_dev))) return rc;
        c.list_host.assign(slots, slots + count);
      }
      std::memset(c.mulbits, 0, sizeof(c.mulbits));
      std::memset(c.snapbits, 0, sizeof(c.snapbits));
      for (size_t i = 0; i < count; ++i) {
        if ((a[i] & ~SLOT_SNAP) != 1u) c.mulbits[i >> 5] |= 1u << (i & 31);
        if (a[i] & SLOT_SNAP) c.snapbits[i >> 5] |= 1u << (i & 31);
      }
      c.bits_count = count;
      SlotScope sc(c, c.list_dev, count);
      rc = annie::square_mul(c, 1u, nullptr, true);
      c.bits_count = 0;
      return rc;
    }
    if ((rc = stage_upload(c, slots, a, count, &d, &da))) return rc;
    SlotScope sc(c, d, count);
    return annie::square_mul(c, 1u, da);
  }
  // Graph path: this step's ids and multipliers into the fixed buffer (ids at
  // [0, count), multipliers at [count, 2 count)), then one replay; the copy
  // and the replay are ordered on the stream after the previous replay.
  hipGraphExec_t exec;
  if ((rc = square_graph(c, count, &exec))) return rc;
  if ((rc = stage_upload(c, slots, a, count, &d, &da, c.graph_buf))) return rc;
  HIPCHK(hipGraphLaunch(exec, c.stream));
  return ANNIE_OK;
}

// Kernel of a squaring run for the batch's plan, or null when the plan has
// none (two-pass lengths, other RNS sizes or digit widths).
typedef void (*RunKernel)(const RunArgs);
static RunKernel run_kernel(const Plan& pl, int* threads) {
  const int qs = pl.cbits() >> 5;
  RunKernel f = nullptr;
#define ANNIE_RUN_CASE(M)                                                        \
  case M:                                                                        \
    *threads = (((1 << M) / NTT_R) >= 64) ? ((1 << M) / NTT_R) : 64;             \
    if (pl.K == 3 && qs == 1) f = k_sqr_run<M, 3, 1>;                            \
    if (pl.K == 4 && qs == 1) f = k_sqr_run<M, 4, 1>;                            \
    if (pl.K == 5 && qs == 2) f = k_sqr_run<M, 5, 2>;                            \
    break;
  switch (pl.LN) {
    ANNIE_RUN_CASE(9)
    ANNIE_RUN_CASE(10)
    ANNIE_RUN_CASE(11)
    ANNIE_RUN_CASE(12)
    ANNIE_RUN_CASE(13)
    default: break;
  }
#undef ANNIE_RUN_CASE
  return f;
}

int annie_hip_batch_run_supported(const annie_hip_t* h) {
  if (!h) return 0;
  int t = 0;
  return (!h->c.slotN.empty() && run_kernel(h->c.pl, &t) != nullptr) ? 1 : 0;
}

int annie_hip_batch_square_run(annie_hip_t* h, const uint32_t* slots, size_t count, const uint32_t* mult,
                               size_t steps) {
  if (!h || !mult) return fail(ANNIE_ERR_ARG, "null argument");
  Ctx& c = h->c;
  if (c.slotN.empty()) return fail(ANNIE_ERR_ARG, "not a batch");
  int rc;
  if ((rc = check_slots(c, slots, count))) return rc;
  if (steps < 1 || steps > ANNIE_HIP_RUN_MAX_STEPS) return fail(ANNIE_ERR_ARG, "steps out of range");
  int threads = 0;
  const RunKernel f = run_kernel(c.pl, &threads);
  if (!f) return fail(ANNIE_ERR_SIZE, "no squaring-run kernel for this plan");
  // Per slot: multipliers in [1, 2^31) (bit 31: snapshot), zeros only as a
  // suffix, a snapshot only on the last squaring of the run.
  for (size_t i = 0; i < count; ++i) {
    const uint32_t* m = mult + i * steps;
    size_t len = 0;
    while (len < steps && m[len] != 0u) ++len;
    for (size_t t = len; t < steps; ++t)
      if (m[t] != 0u) return fail(ANNIE_ERR_ARG, "multiplier after the end of a run");
    for (size_t t = 0; t < len; ++t)
      if ((m[t] & SLOT_SNAP) && t + 1 != len) return fail(ANNIE_ERR_ARG, "snapshot before the end of a run");
  }
  Stage& st = c.run_stage;
  if (!st.dev) {
    st.cap = c.nslots_alloc * (ANNIE_HIP_RUN_MAX_STEPS + 1);
    if (hipMalloc(&st.dev, NSTAGE * st.cap * 4) != hipSuccess) return fail(ANNIE_ERR_NOMEM, "run staging");
    if (hipHostMalloc(&st.host, NSTAGE * st.cap * 4) != hipSuccess) return fail(ANNIE_ERR_NOMEM, "run staging");
    for (int e = 0; e < NSTAGE; ++e) HIPCHK(hipEventCreateWithFlags(&st.ev[e], hipEventDisableTiming));
  }
  const size_t words = count * (steps + 1);
  const int e = st.pos;
  st.pos = (st.pos + 1) % NSTAGE;
  if (st.pending[e]) {
    HIPCHK(hipEventSynchronize(st.ev[e]));
    st.pending[e] = false;
  }
  uint32_t* hbuf = st.host + size_t(e) * st.cap;
  uint32_t* dbuf = st.dev + size_t(e) * st.cap;
  std::memcpy(hbuf, slots, count * 4);
  std::memcpy(hbuf + count, mult, count * steps * 4);
  HIPCHK(hipMemcpyAsync(dbuf, hbuf, words * 4, hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipEventRecord(st.ev[e], c.stream));
  st.pending[e] = true;

  RunArgs R{};
  R.nt = c.nbase;
  R.nt.m = c.pl.LN;
  R.nt.m1 = 0;
  R.ca = c.cbase;
  R.slot_ids = dbuf;
  R.mult = dbuf + count;
  R.slot_stride = c.slot_stride;
  R.steps = int(steps);
  R.r0 = c.reg(R0);
  R.wr = c.reg(RWR);
  R.wq = c.reg(RWQ);
  R.tr = c.reg(RTR);
  R.sq = c.reg(RSQ);
  R.tq = c.tq();
  R.snap = c.reg(RSNAP);
  R.nir = c.reg(RNIR);
  R.nq = c.reg(RNQ);
  R.fwdR = c.fwdR();
  R.invR = c.invR();
  R.fwdQ = c.fwdQ();
  R.invQ = c.invQ();
  const uint32_t S = 1u << c.pl.LN;
  const uint32_t lines = uint32_t(threads) / (S / NTT_R);
  const size_t shm = std::max<size_t>(size_t(lines) * (S + (S >> 5)), size_t(c.pl.K + 1) * (S / CARRY_L)) * 4;
  ProfScope ps(c, "k_sqr_run<" + std::to_string(c.pl.LN) + "," + std::to_string(c.pl.K) + ">");
  hipLaunchKernelGGL(f, dim3(unsigned(count)), dim3(unsigned(threads)), shm, c.stream, R);
  HIPCHK(hipGetLastError());
  return ANNIE_OK;
}

int annie_hip_batch_set_graphs(annie_hip_t* h, int on) {
  if (!h) return fail(ANNIE_ERR_ARG, "null context");
  Ctx& c = h->c;
  if (c.slotN.empty()) return fail(ANNIE_ERR_ARG, "not a batch");
  if (on && !c.graph_buf) {
    if (hipMalloc(&c.graph_buf, 2 * c.nslots_alloc * 4) != hipSuccess) return fail(ANNIE_ERR_NOMEM, "graph buffers");
  }
  HIPCHK(hipStreamSynchronize(c.stream));
  c.use_graphs = on != 0;
  return ANNIE_OK;
}

int annie_hip_batch_mul(annie_hip_t* h, size_t dst, size_t src, const uint32_t* slots, size_t count) {
  if (!h) return fail(ANNIE_ERR_ARG, "null context");
  Ctx& c = h->c;
  if (dst > 2 || src > 3) return fail(ANNIE_ERR_ARG, "register index out of range");
  if (src == 3) src = RSNAP;
  int rc;
  if ((rc = check_slots(c, slots, count))) return rc;
  const uint32_t* d;
  if ((rc = stage_upload(c, slots, nullptr, count, &d, nullptr))) return rc;
  SlotScope sc(c, d, count);
  return annie::mul(c, int(dst), int(src));
}

int annie_hip_batch_get(annie_hip_t* h, size_t reg, const uint32_t* slots, size_t count, uint64_t* limbs,
                        size_t stride, size_t* nout) {
  if (!h || !limbs || !nout) return fail(ANNIE_ERR_ARG, "null argument");
  Ctx& c = h->c;
  if (reg > 2) return fail(ANNIE_ERR_ARG, "register index out of range");
  int rc;
  if ((rc = check_slots(c, slots, count))) return rc;
  const uint32_t* d;
  if ((rc = stage_upload(c, slots, nullptr, count, &d, nullptr))) return rc;
  std::vector<size_t> sl(slots, slots + count);
  std::vector<std::unique_ptr<Mpz>> z(count);
  std::vector<const mpz_t*> mods(count);
  std::vector<mpz_t*> zs(count);
  for (size_t i = 0; i < count; ++i) {
    z[i] = std::make_unique<Mpz>();
    zs[i] = &z[i]->v;
    mods[i] = &c.slotN[slots[i]]->v;
  }
  {
    SlotScope sc(c, d, count);
    if ((rc = get_values(c, int(reg), sl.data(), mods.data(), zs.data(), count))) return rc;
  }
  for (size_t i = 0; i < count; ++i)
    if ((rc = mpz_to_limbs(z[i]->v, limbs + i * stride, stride, nout + i))) return rc;
  return ANNIE_OK;
}

int annie_hip_batch_sync(annie_hip_t* h) {
  if (!h) return fail(ANNIE_ERR_ARG, "null context");
  HIPCHK(hipStreamSynchronize(h->c.stream));
  return ANNIE_OK;
}

void annie_hip_destroy(annie_hip_t* h) { annie::destroy(h); }

const char* annie_hip_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
