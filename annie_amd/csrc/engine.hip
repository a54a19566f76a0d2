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
#include <cstring>
#include <memory>
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
  RTMP,                    // scratch for get()
  NREG
};

struct Ctx {
  Plan pl;
  int device = 0;
  int m1 = 0, m2 = 0;
  size_t n = 0;
  hipStream_t stream = nullptr;
  uint32_t* regs = nullptr;   // NREG * K * n words
  uint32_t* raw = nullptr;    // 2 * DIGIT_LIMBS * n words: T_Q digits, get() digits
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
    (void)hipEventRecord(c.ev[2 * rec], c.stream);
  }
  ~ProfScope() {
    if (rec != size_t(-1)) (void)hipEventRecord(c.ev[2 * rec + 1], c.stream);
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

static int launch_row(Ctx& c, int mode, const NttJob* jobs, int nj) {
  NttArgs A = c.nbase;
  for (int i = 0; i < nj; ++i) A.job[i] = jobs[i];
  A.m = c.m2;
  A.m1 = c.m1;
  const uint32_t S = 1u << c.m2, Tl = S / NTT_R;
  const uint32_t rows = 1u << c.m1;
  uint32_t lines = (Tl >= 64) ? 1u : 64u / Tl;
  if (lines > rows) lines = rows;
  A.cols = int(lines);
  const dim3 grid(rows / lines, c.pl.K, nj), block(lines * Tl);
  const size_t shm = size_t(lines) * (S + (S >> 5)) * sizeof(uint32_t);
  ProfScope ps(c, "k_row<" + std::to_string(c.m2) + "," + row_mode_name(mode) + ">x" + std::to_string(nj));
  switch (c.m2) {
    case 5: launch_row_m<5>(mode, grid, block, shm, c.stream, A); break;
    case 6: launch_row_m<6>(mode, grid, block, shm, c.stream, A); break;
    case 7: launch_row_m<7>(mode, grid, block, shm, c.stream, A); break;
    case 8: launch_row_m<8>(mode, grid, block, shm, c.stream, A); break;
    case 9: launch_row_m<9>(mode, grid, block, shm, c.stream, A); break;
    case 10: launch_row_m<10>(mode, grid, block, shm, c.stream, A); break;
    case 11: launch_row_m<11>(mode, grid, block, shm, c.stream, A); break;
    case 12: launch_row_m<12>(mode, grid, block, shm, c.stream, A); break;
    case 13: launch_row_m<13>(mode, grid, block, shm, c.stream, A); break;
    default: return fail(ANNIE_ERR_SIZE, "no row kernel for this length");
  }
  HIPCHK(hipGetLastError());
  return ANNIE_OK;
}

static int launch_col(Ctx& c, bool inverse, const NttJob* jobs, int nj) {
  NttArgs A = c.nbase;
  for (int i = 0; i < nj; ++i) A.job[i] = jobs[i];
  A.m = c.m1;
  A.m1 = c.m1;
  const uint32_t S = 1u << c.m1, Tl = S / NTT_R;
  const uint32_t n2 = uint32_t(c.n >> c.m1);
  // Up to 16 columns of Tl threads each per block (C * 2^m1 * 4 bytes of LDS).
  uint32_t C = 1024u / Tl;
  if (C < 1) C = 1;
  if (C > 16) C = 16;
  if (C > n2) C = n2;
  A.cols = int(C);
  const dim3 grid(n2 / C, c.pl.K, nj), block(C * Tl);
  const size_t shm = size_t(C) * S * sizeof(uint32_t);
  ProfScope ps(c, std::string(inverse ? "k_col_inv<" : "k_col_fwd<") + std::to_string(c.m1) + ">x" + std::to_string(nj));
  switch (c.m1) {
    case 7: launch_col_m<7>(inverse, grid, block, shm, c.stream, A); break;
    case 8: launch_col_m<8>(inverse, grid, block, shm, c.stream, A); break;
    case 9: launch_col_m<9>(inverse, grid, block, shm, c.stream, A); break;
    case 10: launch_col_m<10>(inverse, grid, block, shm, c.stream, A); break;
    case 11: launch_col_m<11>(inverse, grid, block, shm, c.stream, A); break;
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

static int launch_carry(Ctx& c, const CarryJob* jobs, int nj) {
  CarryArgs A = c.cbase;
  for (int i = 0; i < nj; ++i) A.job[i] = jobs[i];
  const size_t nb = c.n / CARRY_L;
  // CARRY_WG - 1 owned blocks per workgroup (thread 0 only hands over a carry).
  const dim3 grid(unsigned((nb + CARRY_WG - 2) / (CARRY_WG - 1)), nj), block(CARRY_WG);
  const int qs = c.pl.cbits() >> 5;
  if (qs > 2) return fail(ANNIE_ERR_SIZE, "digit width above 95 bits");
  ProfScope ps(c, "k_carry<" + std::to_string(c.pl.K) + ">x" + std::to_string(nj));
  switch (c.pl.K * 4 + qs) {
#define ANNIE_CARRY_CASE(K, QS) \
  case K * 4 + QS: hipLaunchKernelGGL((k_carry<K, QS>), grid, block, 0, c.stream, A); break;
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
  CarryJob j;
  j.src = src;
  j.dst = dst;
  j.sub = sub;
  j.a = a;
  j.nega = nega;
  j.canon = canon;
  j.raw_out = raw_out;
  return j;
}

// reduceMMGW (src/transformGPU.h:523-538): with T_R in RTR and T_Q (raw digits)
// in tq(): m' = T_R N^-1 mod R, canonical, + 2; dst = (m' + 2) N - T_Q mod Q.
static int reduce(Ctx& c, uint32_t* dst, bool raw_out) {
  int rc;
  NttJob j = job(c.reg(RTR), c.reg(RWR), c.fwdR(), c.invR(), c.reg(RNIR));
  if ((rc = convolve(c, ROW_MULC, &j, 1))) return rc;
  CarryJob cj = cjob(c.reg(RWR), c.reg(RTR), 1u, 0, 1, 0);
  if ((rc = launch_carry(c, &cj, 1))) return rc;
  j = job(c.reg(RTR), c.reg(RWQ), c.fwdQ(), c.invQ(), c.reg(RNQ));
  if ((rc = convolve(c, ROW_MULC, &j, 1))) return rc;
  cj = cjob(c.reg(RWQ), dst, 1u, 1, 0, raw_out ? 1 : 0, c.tq());
  return launch_carry(c, &cj, 1);
}

// squareMMGW (src/transformGPU.h:556-568).
static int square_mul(Ctx& c, uint32_t a) {
  int rc;
  const NttJob j[2] = {job(c.reg(R0), c.reg(RWR), c.fwdR(), c.invR()), job(c.reg(R0), c.reg(RWQ), c.fwdQ(), c.invQ())};
  if ((rc = convolve(c, ROW_SQR, j, 2))) return rc;
  const CarryJob cj[2] = {cjob(c.reg(RWR), c.reg(RTR), a, 0, 0, 0), cjob(c.reg(RWQ), c.tq(), a, 1, 0, 1)};
  if ((rc = launch_carry(c, cj, 2))) return rc;
  return reduce(c, c.reg(R0), false);
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
  HIPCHK(hipMemcpyAsync(c.reg(dst), c.reg(src), size_t(c.pl.K) * c.n * 4, hipMemcpyDeviceToDevice, c.stream));
  return ANNIE_OK;
}

// Coefficients of x in base B = 2^cbits (bits [i*cbits, (i+1)*cbits) of x),
// as K residue planes (src/transformGPU.h:323-360).
static void mpz_to_residues(const Ctx& c, const mpz_t x, std::vector<uint32_t>& out) {
  const size_t n = c.n;
  const int cb = c.pl.cbits(), K = c.pl.K;
  out.assign(size_t(K) * n, 0);
  const size_t xs = mpz_size(x);
  const mp_limb_t* xp = mpz_limbs_read(x);
  auto word = [&](size_t pos, int nb) -> uint64_t {  // nb <= 64 bits starting at pos
    const size_t li = pos / 64;
    const int sh = int(pos % 64);
    uint64_t lo = li < xs ? uint64_t(xp[li]) : 0, hi = (li + 1) < xs ? uint64_t(xp[li + 1]) : 0;
    uint64_t r = lo >> sh;
    if (sh) r |= hi << (64 - sh);
    return nb == 64 ? r : (r & ((uint64_t(1) << nb) - 1));
  };
  for (size_t i = 0; i < n; ++i) {
    const size_t pos = i * size_t(cb);
    const int b0 = cb < 64 ? cb : 64;
    const uint64_t lo = word(pos, b0);
    const uint64_t hi = cb > 64 ? word(pos + 64, cb - 64) : 0;
    for (int k = 0; k < K; ++k) {
      const uint32_t p = PRIMES[k];
      // (hi * 2^64 + lo) mod p
      const uint64_t t64 = uint64_t(pow2mod(64, p));
      const uint64_t r = ((hi % p) * t64 + (lo % p)) % p;
      out[size_t(k) * n + i] = uint32_t(r);
    }
  }
}

// Signed digits (DIGIT_LIMBS planes, two's complement) -> integer sum d_i B^i.
static void digits_to_mpz(const Ctx& c, const std::vector<uint32_t>& d, mpz_t z) {
  const size_t n = c.n;
  const int cb = c.pl.cbits();
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

static int upload_mpz(Ctx& c, int r, const mpz_t x) {
  std::vector<uint32_t> h;
  mpz_to_residues(c, x, h);
  HIPCHK(hipMemcpyAsync(c.reg(r), h.data(), h.size() * 4, hipMemcpyHostToDevice, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  return ANNIE_OK;
}

static int scale_reg(Ctx& c, int r) {
  uint32_t s[KMAX] = {0, 0, 0, 0, 0};
  for (int k = 0; k < c.pl.K; ++k) s[k] = c.pc.sq_scale[k];  // mont(x, n^-1 2^64) = x n^-1 2^32
  const dim3 grid(unsigned((c.n + 255) / 256), c.pl.K), block(256);
  hipLaunchKernelGGL(k_scale, grid, block, 0, c.stream, c.reg(r), c.n, c.pl.K, c.nbase, s[0], s[1], s[2], s[3], s[4]);
  HIPCHK(hipGetLastError());
  return ANNIE_OK;
}

static int get_value(Ctx& c, int src, mpz_t z) {
  int rc;
  // T_Q = src as raw digits (a carry pass over its own residues), T_R = src.
  const CarryJob cj = cjob(c.reg(src), c.tq(), 1u, 1, 0, 1);
  if ((rc = launch_carry(c, &cj, 1))) return rc;
  if ((rc = copy_reg(c, RTR, src))) return rc;
  if ((rc = reduce(c, c.rawtmp(), true))) return rc;
  std::vector<uint32_t> d(size_t(DIGIT_LIMBS) * c.n);
  HIPCHK(hipMemcpyAsync(d.data(), c.rawtmp(), d.size() * 4, hipMemcpyDeviceToHost, c.stream));
  HIPCHK(hipStreamSynchronize(c.stream));
  digits_to_mpz(c, d, z);
  mpz_mod(z, z, c.N);
  return ANNIE_OK;
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
  if (c.regs) (void)hipFree(c.regs);
  if (c.raw) (void)hipFree(c.raw);
  if (c.tw) (void)hipFree(c.tw);
  for (hipEvent_t e : c.ev) (void)hipEventDestroy(e);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  mpz_clears(c.N, c.factor, nullptr);
  delete h;
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
  if (c.pl.LN < 5) return fail(ANNIE_ERR_SIZE, "NTT size < 2^5 is not supported by this engine");
  c.n = c.pl.n();
  c.m1 = (c.pl.LN <= 13) ? 0 : c.pl.LN / 2;
  c.m2 = c.pl.LN - c.m1;
  c.device = device;

  // R = B^n - 1; N must be invertible mod R (src/transformGPU.h:652-661).
  mpz_t R, t;
  mpz_inits(R, t, nullptr);
  mpz_set_ui(R, 1);
  mpz_mul_2exp(R, R, mp_bitcnt_t(c.pl.cbits()) << c.pl.LN);
  mpz_sub_ui(R, R, 1);
  struct Clear {
    mpz_t& a;
    mpz_t& b;
    ~Clear() { mpz_clears(a, b, nullptr); }
  } clear_rt{R, t};
  if (mpz_invert(t, c.N, R) == 0) {
    mpz_gcd(c.factor, c.N, R);
    c.init_failed = true;
    *out = h.release();
    return ANNIE_INIT_FAILED;
  }

  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
  const size_t K = size_t(c.pl.K), n = c.n;
  if (hipMalloc(&c.regs, size_t(NREG) * K * n * 4) != hipSuccess) return fail(ANNIE_ERR_NOMEM, "register file");
  if (hipMalloc(&c.raw, 2 * size_t(DIGIT_LIMBS) * n * 4) != hipSuccess) return fail(ANNIE_ERR_NOMEM, "digit buffers");
  if (hipMalloc(&c.tw, 4 * K * n * sizeof(TwDev)) != hipSuccess) return fail(ANNIE_ERR_NOMEM, "twiddles");
  HIPCHK(hipMemsetAsync(c.regs, 0, size_t(NREG) * K * n * 4, c.stream));

  c.pc = make_consts(c.pl);
  for (int k = 0; k < KMAX; ++k) {
    c.nbase.p[k] = c.cbase.p[k] = c.pc.p[k];
    c.nbase.q[k] = c.cbase.q[k] = c.pc.q[k];
    c.nbase.sc[k] = c.pc.sq_scale[k];
    c.nbase.scq[k] = c.pc.sq_scale_q[k];
    c.cbase.c32[k] = c.pc.c32[k];
    c.cbase.c32q[k] = c.pc.c32q[k];
    c.cbase.c64[k] = c.pc.c64[k];
    c.cbase.c64q[k] = c.pc.c64q[k];
    c.cbase.c32n[k] = c.pc.p[k] - c.pc.c32[k];
    c.cbase.c32nq[k] = c.cbase.c32n[k] * c.pc.q[k];
    for (int j = 0; j < KMAX; ++j) {
      c.cbase.ginv[k][j] = c.pc.ginv[k][j];
      c.cbase.ginvq[k][j] = c.pc.ginvq[k][j];
    }
  }
  c.nbase.LN = c.cbase.LN = c.pl.LN;
  c.cbase.cbits = c.pl.cbits();

  // Twiddle tables.
  {
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
  }

  int rc;
  // N_Q: negacyclic transform of N; Ninv_R: cyclic transform of N^-1 mod R
  // (src/transformGPU.h:737-741), both scaled by n^-1 2^32 for the pointwise step.
  if ((rc = upload_mpz(c, RTMP, c.N))) return rc;
  {
    NttJob j = job(c.reg(RTMP), c.reg(RNQ), c.fwdQ(), c.invQ());
    if ((rc = forward(c, &j, 1))) return rc;
    if ((rc = scale_reg(c, RNQ))) return rc;
  }
  if ((rc = upload_mpz(c, RTMP, t))) return rc;
  {
    NttJob j = job(c.reg(RTMP), c.reg(RNIR), c.fwdR(), c.invR());
    if ((rc = forward(c, &j, 1))) return rc;
    if ((rc = scale_reg(c, RNIR))) return rc;
  }
  // The MMGW form of 1 is R/2 mod N (src/transformGPU.h:743-747).
  mpz_mod(t, R, c.N);
  if (mpz_odd_p(t)) mpz_add(t, t, c.N);
  mpz_fdiv_q_2exp(t, t, 1);
  if ((rc = upload_mpz(c, RONE, t))) return rc;
  HIPCHK(hipStreamSynchronize(c.stream));
  *out = h.release();
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

void annie_hip_destroy(annie_hip_t* h) { annie::destroy(h); }

const char* annie_hip_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
