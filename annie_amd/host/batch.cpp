// Batched and served tests: many Fermat / Euler (Proth) tests run together,
// one engine batch per transform-length class, with continuous admission.
//
// Each test runs the protocol of annie::check (src/annie.h:293-404): set1,
// the hot loop of square_mul(bit ? a : 1) with the Gerbicz-Li product
// mul(Reg1, Reg0) every B iterations (src/annie.h:319-341), the final
// d(t+1) = d(t) * result, the residue and verdict (src/annie.h:343-366) and
// the Gerbicz-Li verification a^(res - 2^B) ... (src/annie.h:368-404).  Here
// that sequence is a state machine per test that yields one register
// operation at a time; every tick the driver takes one operation from each
// test of a class and issues them grouped by kind, so one set of launches
// squares every test of the class.  A test that finishes frees its slot and
// the next waiting test of its class is admitted at once (continuous
// batching); tests that arrive later (serving) wait until their arrival time.
// Classes are keyed by the plan (K, pck, LN) the reference's size selection
// gives for (N, a), src/transformGPU.h:798-857, and the digit-width class of
// its lbase (pck lbase >> 5): plans that differ only in lbase share a batch.
#include <gmp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <mutex>
#include <atomic>
#include <tuple>
#include <vector>

#include "annie_hip.h"
#include "annie_host.h"
#include "transform.h"

namespace annie_host {

namespace {

struct Big {
  mpz_t v;
  Big() { mpz_init(v); }
  ~Big() { mpz_clear(v); }
  Big(const Big&) = delete;
  Big& operator=(const Big&) = delete;
};

// K, pck, LN and the digit-width class pck lbase >> 5: moduli whose plans
// differ only in lbase within a class share one engine batch.
typedef std::tuple<int, int, int, int> PlanKey;

// A batch of slots of one plan.  Operations act on the listed slots.
class BatchEngine {
 public:
  virtual ~BatchEngine() {}
  // 0: loaded; 1: gcd(N, R) > 1, factor set (src/transformGPU.h:652-661).
  virtual int load(size_t slot, const mpz_t N, int32_t a, mpz_t factor) = 0;
  virtual void set1(const std::vector<uint32_t>& s) = 0;
  virtual void copy(int dst, int src, const std::vector<uint32_t>& s) = 0;
  virtual void square_mul(const std::vector<uint32_t>& s, const std::vector<uint32_t>& a) = 0;
  virtual void mul(int dst, int src, const std::vector<uint32_t>& s) = 0;
  virtual void get(int reg, const std::vector<uint32_t>& s, std::vector<Big*>& out) = 0;
  // Squaring runs: slot s[i] takes the multipliers mult[i * steps ...] up to
  // its first zero (annie_hip_batch_square_run).
  virtual bool can_run() const { return false; }
  virtual void square_run(const std::vector<uint32_t>& s, const std::vector<uint32_t>& mult, size_t steps) = 0;
};

// The gfx950 engine's batch C ABI.
class HipBatch : public BatchEngine {
  annie_hip_t* h_ = nullptr;
  size_t stride_ = 0;
  std::vector<uint64_t> limbs_;
  std::vector<size_t> nout_;

  void check(int rc, const char* what) {
    if (rc != ANNIE_OK) throw std::runtime_error(std::string(what) + ": " + annie_hip_last_error());
  }

 public:
  // lbase: that of one of the class's moduli.
  HipBatch(const PlanKey& k, int lbase, size_t slots, int device, bool graphs, int row_lines, int col_cols) {
    const int K = std::get<0>(k), pck = std::get<1>(k), LN = std::get<2>(k);
    check(annie_hip_batch_create(K, pck, LN, lbase, slots, device, &h_), "batch create");
    if (graphs) check(annie_hip_batch_set_graphs(h_, 1), "graphs");
    if ((row_lines != 0 || col_cols != 0) && annie_hip_schedule_fits(LN, row_lines, col_cols))
      check(annie_hip_set_schedule(h_, row_lines, col_cols), "schedule");
    // limbs of a value below B^n with the widest digit of the class
    stride_ = ((size_t(32) * size_t(std::get<3>(k) + 1)) << LN) / 64 + 2;
  }
  ~HipBatch() override { annie_hip_destroy(h_); }
  int load(size_t slot, const mpz_t N, int32_t a, mpz_t factor) override {
    const size_t nl = mpz_size(N);
    std::vector<uint64_t> l(nl);
    mpz_export(l.data(), nullptr, -1, 8, 0, 0, N);
    std::vector<uint64_t> f(nl + 1);
    size_t nf = 0;
    const int rc = annie_hip_batch_load(h_, slot, l.data(), nl, a, f.data(), f.size(), &nf);
    if (rc == ANNIE_INIT_FAILED) {
      mpz_import(factor, nf, -1, 8, 0, 0, f.data());
      return 1;
    }
    check(rc, "batch load");
    return 0;
  }
  void set1(const std::vector<uint32_t>& s) override { check(annie_hip_batch_set1(h_, s.data(), s.size()), "set1"); }
  void copy(int dst, int src, const std::vector<uint32_t>& s) override {
    check(annie_hip_batch_copy(h_, size_t(dst), size_t(src), s.data(), s.size()), "copy");
  }
  void square_mul(const std::vector<uint32_t>& s, const std::vector<uint32_t>& a) override {
    check(annie_hip_batch_square_mul(h_, s.data(), a.data(), s.size()), "square_mul");
  }
  void mul(int dst, int src, const std::vector<uint32_t>& s) override {
    check(annie_hip_batch_mul(h_, size_t(dst), size_t(src), s.data(), s.size()), "mul");
  }
  bool can_run() const override { return annie_hip_batch_run_supported(h_) != 0; }
  void square_run(const std::vector<uint32_t>& s, const std::vector<uint32_t>& mult, size_t steps) override {
    check(annie_hip_batch_square_run(h_, s.data(), s.size(), mult.data(), steps), "square_run");
  }
  void get(int reg, const std::vector<uint32_t>& s, std::vector<Big*>& out) override {
    limbs_.assign(stride_ * s.size(), 0);
    nout_.assign(s.size(), 0);
    check(annie_hip_batch_get(h_, size_t(reg), s.data(), s.size(), limbs_.data(), stride_, nout_.data()), "get");
    for (size_t i = 0; i < s.size(); ++i) mpz_import(out[i]->v, nout_[i], -1, 8, 0, 0, limbs_.data() + i * stride_);
  }
};

// The same operations on exact GMP transforms, one per slot (CPU; used to
// test the driver and the state machine without a GPU).
class GmpBatch : public BatchEngine {
  std::vector<std::unique_ptr<transform>> t_;

 public:
  explicit GmpBatch(size_t slots) : t_(slots) {}
  int load(size_t slot, const mpz_t N, int32_t a, mpz_t factor) override {
    t_[slot].reset(create_transform_gmp(N, a));
    if (t_[slot]->init_failed()) {
      t_[slot]->get_factor(factor);
      t_[slot].reset();
      return 1;
    }
    return 0;
  }
  void set1(const std::vector<uint32_t>& s) override {
    for (uint32_t i : s) t_[i]->set1();
  }
  void copy(int dst, int src, const std::vector<uint32_t>& s) override {
    for (uint32_t i : s) t_[i]->copy(size_t(dst), size_t(src));
  }
  // A snapshot square (bit 31 of a) takes the Gerbicz-Li product at once; the
  // later product with the snapshot register is then a no-op (same result).
  void square_mul(const std::vector<uint32_t>& s, const std::vector<uint32_t>& a) override {
    for (size_t j = 0; j < s.size(); ++j) {
      t_[s[j]]->square_mul(int32_t(a[j] & 0x7fffffffu));
      if (a[j] & 0x80000000u) t_[s[j]]->mul(1, 0);
    }
  }
  void mul(int dst, int src, const std::vector<uint32_t>& s) override {
    if (src == 3) return;
    for (uint32_t i : s) t_[i]->mul(size_t(dst), size_t(src));
  }
  void get(int reg, const std::vector<uint32_t>& s, std::vector<Big*>& out) override {
    for (size_t j = 0; j < s.size(); ++j) t_[s[j]]->get(out[j]->v, size_t(reg));
  }
  bool can_run() const override { return true; }
  void square_run(const std::vector<uint32_t>& s, const std::vector<uint32_t>& mult, size_t steps) override {
    for (size_t j = 0; j < s.size(); ++j)
      for (size_t t = 0; t < steps && mult[j * steps + t] != 0u; ++t) square_mul({s[j]}, {mult[j * steps + t]});
  }
};

// Exponent block of the Gerbicz-Li proof (src/annie.h:95-101).
int gl_block_size(size_t size) {
  int lg = -1;
  for (uint32_t v = uint32_t(size); v; v >>= 1) ++lg;
  const size_t L = size_t(2) << (lg / 2);
  return int((size - 1) / L) + 1;
}

enum OpKind { OP_SET1, OP_COPY, OP_SQR, OP_MUL, OP_GET, OP_DONE };
struct Op {
  OpKind kind;
  int dst, src;
  uint32_t a;
};

// Multiplier bit asking the engine to keep a copy of the new r0 (the snapshot
// register, 3) for a Gerbicz-Li product taken later.
constexpr uint32_t SNAP = 0x80000000u;

// The protocol of one test as a sequence of register operations.  The
// Gerbicz-Li product d(t) *= r0 of a loop step (src/annie.h:338) is deferred:
// that step's square_mul also writes r0 to the snapshot register, and
// mul(1, 3) follows at a tick where the driver flushes pending products for
// all tests together, or before this test's next snapshot or the end of its
// loop.  The products commute, so d(t) is unchanged.
class TestMachine {
  enum Phase { SET1, COPY1, LOOP, FIN_COPY, FIN_MUL, GET_T, EULER_SQR, GET_RES, VSET1, V1, VMUL, V2, GET_R1, GET_R2, DONE };
  Phase ph_ = SET1;
  long i_ = 0;
  int B_ = 1, rs_ = 0;
  bool pending_ = false;
  Op last_{OP_DONE, 0, 0, 0};
  Big e_, res_, r1_;
  const __mpz_struct* N_;

  bool boundary() const { return (i_ % B_ == 0) && (i_ / B_ != 0); }

 public:
  int32_t a;
  bool euler;
  long iterations = 0;
  bool is_prp = false, gl_ok = false;
  uint64_t res64 = 0;

  TestMachine(const mpz_t N, int32_t base, bool is_euler) : N_(N), a(base), euler(is_euler) {
    mpz_sub_ui(e_.v, N, 1);
    if (euler) mpz_tdiv_q_2exp(e_.v, e_.v, 1);
    B_ = gl_block_size(mpz_sizeinbase(e_.v, 2));
    i_ = long(mpz_sizeinbase(e_.v, 2)) - 1;
  }
  bool done() const { return ph_ == DONE; }
  bool pending() const { return pending_; }

  // Next operation; `flush`: take a pending Gerbicz-Li product now.
  Op op(bool flush) {
    last_ = next(flush);
    return last_;
  }

  Op next(bool flush) const {
    switch (ph_) {
      case SET1: return {OP_SET1, 0, 0, 0};
      case COPY1: return {OP_COPY, 1, 0, 0};
      case LOOP:
        if (pending_ && (flush || boundary())) return {OP_MUL, 1, 3, 0};
        return {OP_SQR, 0, 0, (mpz_tstbit(e_.v, mp_bitcnt_t(i_)) ? uint32_t(a) : 1u) | (boundary() ? SNAP : 0u)};
      case FIN_COPY:
        if (pending_) return {OP_MUL, 1, 3, 0};
        return {OP_COPY, 2, 0, 0};
      case FIN_MUL: return {OP_MUL, 2, 1, 0};
      case GET_T: return {OP_GET, 0, 0, 0};
      case EULER_SQR: return {OP_SQR, 0, 0, 1u};
      case GET_RES: return {OP_GET, 0, 0, 0};
      case VSET1: return {OP_SET1, 0, 0, 0};
      case V1: return {OP_SQR, 0, 0, mpz_tstbit(res_.v, mp_bitcnt_t(i_)) ? uint32_t(a) : 1u};
      case VMUL: return {OP_MUL, 0, 1, 0};
      case V2: return {OP_SQR, 0, 0, (i_ < rs_ && mpz_tstbit(res_.v, mp_bitcnt_t(i_))) ? uint32_t(a) : 1u};
      case GET_R1: return {OP_GET, 0, 0, 0};
      case GET_R2: return {OP_GET, 2, 0, 0};
      default: return {OP_DONE, 0, 0, 0};
    }
  }

  // Advances past the operation last returned by op(); v is the value of a GET.
  void advance(const mpz_t v) {
    if (last_.kind == OP_MUL && last_.src == 3) {  // deferred Gerbicz-Li product taken
      pending_ = false;
      return;
    }
    switch (ph_) {
      case SET1: ph_ = COPY1; break;
      case COPY1: ph_ = LOOP; break;
      case LOOP:
        ++iterations;
        if (last_.a & SNAP) pending_ = true;
        if (--i_ < 0) ph_ = FIN_COPY;
        break;
      case FIN_COPY: ph_ = FIN_MUL; break;
      case FIN_MUL: ph_ = GET_T; break;
      case GET_T: {  // result ?= -1 (Euler) or 1 (Fermat), src/annie.h:349-356
        Big t;
        if (euler) {
          mpz_add_ui(t.v, v, 1);
          if (mpz_cmp(t.v, N_) >= 0) mpz_sub(t.v, t.v, N_);
        } else {
          mpz_sub_ui(t.v, v, 1);
        }
        is_prp = mpz_sgn(t.v) == 0;
        ph_ = euler ? EULER_SQR : GET_RES;
        break;
      }
      case EULER_SQR: ph_ = GET_RES; break;
      case GET_RES: {
        res64 = mpz_size(v) ? uint64_t(mpz_getlimbn(v, 0)) : 0;
        // res = sum of the B-bit chunks of e (src/annie.h:368-375)
        Big t;
        mpz_set_ui(res_.v, 0);
        while (mpz_sgn(e_.v) != 0) {
          mpz_tdiv_r_2exp(t.v, e_.v, mp_bitcnt_t(B_));
          mpz_add(res_.v, res_.v, t.v);
          mpz_tdiv_q_2exp(e_.v, e_.v, mp_bitcnt_t(B_));
        }
        rs_ = int(mpz_sizeinbase(res_.v, 2));
        ph_ = VSET1;
        break;
      }
      case VSET1:
        i_ = rs_ - 1;
        ph_ = (i_ >= B_) ? V1 : VMUL;
        break;
      case V1:
        if (--i_ < B_) ph_ = VMUL;
        break;
      case VMUL:
        i_ = B_ - 1;
        ph_ = V2;
        break;
      case V2:
        if (--i_ < 0) ph_ = GET_R1;
        break;
      case GET_R1:
        mpz_set(r1_.v, v);
        ph_ = GET_R2;
        break;
      case GET_R2:
        gl_ok = mpz_cmp(r1_.v, v) == 0;
        ph_ = DONE;
        break;
      default: break;
    }
  }
};

std::string hex64(uint64_t u) {
  char b[32];
  std::snprintf(b, sizeof(b), "%016llX", (unsigned long long)u);
  return b;
}

std::string short_name(const mpz_t n) {
  std::string s(mpz_sizeinbase(n, 10) + 2, '\0');
  mpz_get_str(&s[0], 10, n);
  s = s.c_str();
  if (s.size() <= 40) return s;
  return s.substr(0, 8) + "[" + std::to_string(s.size() - 16) + " other digits]" + s.substr(s.size() - 8);
}

struct Job {
  size_t idx;
  Big N;
  int32_t a = 2;
  bool euler = false;
  double arrival = 0;
  PlanKey key;
  std::unique_ptr<TestMachine> m;
};

struct Class {
  PlanKey key;
  int lbase = 0;               // lbase of one of the class's moduli
  std::unique_ptr<BatchEngine> eng;
  size_t slots = 0;
  std::vector<Job*> slot;      // occupant per slot (null: free)
  std::deque<Job*> queue;      // waiting tests, in arrival order
  size_t active = 0;
  size_t remaining = 0;        // tests of this class not yet finished
  uint64_t ticks = 0;
};

constexpr uint64_t FLUSH = 16;

}  // namespace

static int batch_run(const annie_batch_test* tests, size_t count, const annie_batch_opts& o,
                     annie_batch_result* out, annie_batch_stats* st) {
  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  auto now = [&]() { return std::chrono::duration<double>(clock::now() - t0).count(); };
  std::memset(out, 0, count * sizeof(annie_batch_result));
  if (st) std::memset(st, 0, sizeof(*st));

  std::vector<std::unique_ptr<Job>> jobs;
  std::map<PlanKey, Class> classes;
  for (size_t t = 0; t < count; ++t) {
    annie_batch_result& r = out[t];
    r.arrival = tests[t].arrival;
    auto j = std::make_unique<Job>();
    j->idx = t;
    mpz_import(j->N.v, tests[t].nlimbs, -1, 8, 0, 0, tests[t].N);
    j->euler = tests[t].is_euler != 0;
    j->arrival = tests[t].arrival;
    // Decided without the loop (src/annie.h:222-247).  The decimal string is
    // only formed for numbers that small (or even): at 2e7 bits it takes
    // seconds, and the messages of tests that run form it when they finish.
    const bool tiny = mpz_sizeinbase(j->N.v, 10) <= 16 && short_name(j->N.v).size() <= 15;
    if (tiny) {
      const std::string ns = short_name(j->N.v);
      const int pp = mpz_probab_prime_p(j->N.v, 20);
      r.status = 4;
      r.is_prp = pp != 0;
      std::snprintf(r.message, sizeof(r.message), "%s", (ns + (pp == 0 ? " is composite." : " is prime.")).c_str());
      continue;
    }
    if (mpz_even_p(j->N.v)) {
      r.status = 4;
      std::snprintf(r.message, sizeof(r.message), "%s is even.", short_name(j->N.v).c_str());
      continue;
    }
    int32_t a = 2;  // src/annie.h:249-254
    if (j->euler) {
      while (mpz_si_kronecker(a, j->N.v) != -1) {
        ++a;
        const int32_t s = int32_t(std::lround(std::sqrt(double(a))));
        if (s * s == a) ++a;
      }
    } else if (tests[t].a_ref != 0) {
      a = tests[t].a_ref;
    }
    j->a = a;
    r.a = a;
    int K = 0, pck = 0, LN = 0, lbase = 0;
    if (annie_hip_select_plan(tests[t].N, tests[t].nlimbs, a, &K, &pck, &LN, &lbase) != ANNIE_OK) {
      r.status = -1;
      std::snprintf(r.message, sizeof(r.message), "Error: %s.", annie_hip_last_error());
      continue;
    }
    if (LN < 4 || LN > 23) {
      r.status = -1;
      std::snprintf(r.message, sizeof(r.message), "Error: no transform of this size.");
      continue;
    }
    j->key = PlanKey(K, pck, LN, (pck * lbase) >> 5);
    Class& c = classes[j->key];
    c.key = j->key;
    c.lbase = lbase;
    c.remaining++;
    jobs.push_back(std::move(j));
  }
  for (auto& j : jobs) classes[j->key].queue.push_back(j.get());
  for (auto& kv : classes)
    std::stable_sort(kv.second.queue.begin(), kv.second.queue.end(),
                     [](const Job* x, const Job* y) { return x->arrival < y->arrival; });

  // Slots per class: every test of the class, capped by max_slots and by an
  // equal share of the HBM budget (the engine's bytes per slot and per batch).
  const double budget = o.hbm_budget_mb > 0 ? o.hbm_budget_mb * 1048576.0 : 64.0 * 1073741824.0;
  const size_t ncls = std::max<size_t>(classes.size(), 1);
  for (auto& kv : classes) {
    Class& c = kv.second;
    size_t per_slot = 0, fixed_bytes = 0;
    if (annie_hip_batch_bytes(std::get<0>(c.key), std::get<2>(c.key), &per_slot, &fixed_bytes) != ANNIE_OK)
      throw std::runtime_error(annie_hip_last_error());
    const double slot_bytes = double(per_slot), fixed = double(fixed_bytes);
    size_t s = c.remaining;
    if (o.max_slots > 0) s = std::min(s, o.max_slots);
    const double share = budget / double(ncls) - fixed;
    s = std::min(s, size_t(std::max(1.0, std::floor(share / slot_bytes))));
    c.slots = std::max<size_t>(s, 1);
  }

  // Host workers: each drives its own classes (class i to worker i mod W,
  // W = min(classes, host_threads)).  A class has its own batch and stream,
  // so a worker waits only for its own classes' work (a get() synchronises
  // one stream), and the launches of different classes run side by side on
  // the GPU instead of taking turns behind one host thread.
  const uint64_t flush_ticks = o.flush_ticks > 0 ? uint64_t(o.flush_ticks) : FLUSH;
  std::vector<Class*> order;
  for (auto& kv : classes) order.push_back(&kv.second);
  // Graph capture (graphs on) stays on one host thread: the runtime rejects
  // allocation and synchronisation calls of other threads while a stream is
  // being captured (seen as a failed test with several workers).
  const size_t want_threads = o.graphs ? 1 : (o.host_threads > 0 ? size_t(o.host_threads) : 16);
  const size_t nthreads = std::max<size_t>(1, std::min(order.size(), want_threads));
  std::mutex io_mu;
  std::atomic<bool> failed{false};
  std::string fail_msg;
  std::vector<double> squarings_w(nthreads, 0.0), setup_w(nthreads, 0.0);

  const double t_start = now();  // time_limit counts from here, set-up excluded
  auto worker = [&](size_t w) {
    std::vector<Class*> mine;
    for (size_t i = w; i < order.size(); i += nthreads) mine.push_back(order[i]);
    double& squarings = squarings_w[w];
    double& setup = setup_w[w];
    std::vector<Big*> vals;
    std::vector<std::unique_ptr<Big>> store;
    // Per-tick operation lists, reused across ticks and classes (cleared, not
    // reallocated, every tick).
    std::vector<uint32_t> set1, sq, sqa, run_ids, run_mult;
    std::map<std::pair<int, int>, std::vector<uint32_t>> cp, ml;
    std::map<int, std::vector<uint32_t>> gt;
    std::vector<uint8_t> handled;
    Big none;
    auto finish = [&](Class& c, size_t s) {
      Job* j = c.slot[s];
      annie_batch_result& r = out[j->idx];
      TestMachine& m = *j->m;
      r.finished = now();
      r.iterations = m.iterations;
      r.is_prp = m.is_prp;
      r.res64 = m.res64;
      r.gl_ok = m.gl_ok;
      const std::string ns = short_name(j->N.v);
      std::string msg;
      if (m.gl_ok) {
        r.status = 0;
        msg = m.is_prp ? ns + " is a " + (m.euler ? "Euler" : "Fermat") + " probable prime, a = " + std::to_string(m.a) +
                             "."
                       : ns + " is composite, res64 = " + hex64(m.res64) + ", a = " + std::to_string(m.a) + ".";
      } else {
        r.status = 1;
        msg = ns + ": validation failed!";
      }
      std::snprintf(r.message, sizeof(r.message), "%s", msg.c_str());
      if (o.verbose) {
        std::lock_guard<std::mutex> lock(io_mu);
        std::fprintf(stderr, "[%.3f s] %s\n", r.finished, msg.c_str());
      }
      c.slot[s] = nullptr;
      c.active--;
      c.remaining--;
      j->m.reset();
    };

    while (!failed.load(std::memory_order_relaxed)) {
      const double tnow = now();
      if (o.time_limit > 0 && tnow - t_start - setup >= o.time_limit) break;
      // Admission: free slots take arrived tests of their class.
      for (Class* cl : mine) {
        Class& c = *cl;
        while (!c.queue.empty() && (!o.realtime || c.queue.front()->arrival <= tnow)) {
          const double ts = now();
          if (!c.eng) {
            if (o.engine == 1) c.eng.reset(new GmpBatch(c.slots));
            else c.eng.reset(new HipBatch(c.key, c.lbase, c.slots, o.device, o.graphs != 0, o.row_lines, o.col_cols));
            c.slot.assign(c.slots, nullptr);
          }
          size_t s = 0;
          while (s < c.slots && c.slot[s]) ++s;
          if (s == c.slots) break;
          Job* j = c.queue.front();
          c.queue.pop_front();
          annie_batch_result& r = out[j->idx];
          r.admitted = now();
          Big f;
          const int lr = c.eng->load(s, j->N.v, j->a, f.v);
          setup += now() - ts;
          if (lr == 1) {  // src/annie.h:269-283
            r.status = 3;
            r.finished = now();
            std::string fs(mpz_sizeinbase(f.v, 10) + 2, '\0');
            mpz_get_str(&fs[0], 10, f.v);
            std::snprintf(r.message, sizeof(r.message), "%s divides %s.", fs.c_str(), short_name(j->N.v).c_str());
            c.remaining--;
            continue;
          }
          j->m = std::make_unique<TestMachine>(j->N.v, j->a, j->euler);
          c.slot[s] = j;
          c.active++;
        }
      }
      // One tick per class with active tests.
      bool any = false;
      for (Class* cl : mine) {
        Class& c = *cl;
        if (c.active == 0) continue;
        any = true;
        set1.clear();
        sq.clear();
        sqa.clear();
        for (auto& g : cp) g.second.clear();
        for (auto& g : ml) g.second.clear();
        for (auto& g : gt) g.second.clear();
        // Deferred Gerbicz-Li products of all tests are taken together every
        // FLUSH ticks (a test forces its own earlier if it must).  In run
        // mode a tick is many squarings long: pending products are taken at
        // every tick, and each test whose next operation is a squaring takes
        // its following squarings (up to run_steps, ending at one that takes
        // a Gerbicz-Li snapshot or before any other operation) in one launch.
        const bool run_mode = o.run_steps > 0 && c.active >= std::max<size_t>(size_t(o.run_min_active), 1) &&
                              c.eng->can_run();
        const bool flush = run_mode || (++c.ticks % flush_ticks) == 0;
        const size_t steps = size_t(std::max(o.run_steps, 1));
        run_ids.clear();
        run_mult.clear();
        handled.assign(c.slots, 0);
        for (size_t s = 0; s < c.slots; ++s) {
          if (!c.slot[s]) continue;
          TestMachine& m = *c.slot[s]->m;
          const Op op = m.op(flush);
          const uint32_t id = uint32_t(s);
          if (run_mode && op.kind == OP_SQR) {
            const size_t row = run_mult.size();
            run_ids.push_back(id);
            run_mult.resize(row + steps, 0u);
            run_mult[row] = op.a;
            m.advance(none.v);
            size_t k = 1;
            if (!(op.a & SNAP)) {
              while (k < steps) {
                const Op nx = m.op(false);
                if (nx.kind != OP_SQR) break;
                run_mult[row + k++] = nx.a;
                m.advance(none.v);
                if (nx.a & SNAP) break;
              }
            }
            squarings += double(k);
            handled[s] = 1;
            continue;
          }
          switch (op.kind) {
            case OP_SET1: set1.push_back(id); break;
            case OP_COPY: cp[{op.dst, op.src}].push_back(id); break;
            case OP_SQR:
              sq.push_back(id);
              sqa.push_back(op.a);
              break;
            case OP_MUL: ml[{op.dst, op.src}].push_back(id); break;
            case OP_GET: gt[op.dst].push_back(id); break;
            default: break;
          }
        }
        if (!set1.empty()) c.eng->set1(set1);
        for (auto& g : cp)
          if (!g.second.empty()) c.eng->copy(g.first.first, g.first.second, g.second);
        if (!sq.empty()) {
          c.eng->square_mul(sq, sqa);
          squarings += double(sq.size());
        }
        if (!run_ids.empty()) c.eng->square_run(run_ids, run_mult, steps);
        for (auto& g : ml)
          if (!g.second.empty()) c.eng->mul(g.first.first, g.first.second, g.second);
        for (auto& g : gt) {
          if (g.second.empty()) continue;
          store.clear();
          vals.clear();
          for (size_t i = 0; i < g.second.size(); ++i) {
            store.push_back(std::make_unique<Big>());
            vals.push_back(store.back().get());
          }
          c.eng->get(g.first, g.second, vals);
          for (size_t i = 0; i < g.second.size(); ++i) {
            handled[g.second[i]] = 1;
            c.slot[g.second[i]]->m->advance(vals[i]->v);
            if (c.slot[g.second[i]]->m->done()) finish(c, g.second[i]);
          }
        }
        for (size_t s = 0; s < c.slots; ++s) {
          if (!c.slot[s] || handled[s]) continue;
          c.slot[s]->m->advance(none.v);
          if (c.slot[s]->m->done()) finish(c, s);
        }
      }
      if (!any) {
        // Nothing running: wait for the next arrival of one of these classes.
        double next = -1;
        for (Class* cl : mine)
          if (!cl->queue.empty())
            next = (next < 0) ? cl->queue.front()->arrival : std::min(next, cl->queue.front()->arrival);
        if (next < 0) break;
        const double wait = next - now();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
      }
    }
  };
  auto guarded = [&](size_t w) {
    try {
      worker(w);
    } catch (const std::exception& ex) {
      std::lock_guard<std::mutex> lock(io_mu);
      if (!failed.exchange(true)) fail_msg = ex.what();
    }
  };
  std::vector<std::thread> pool;
  for (size_t w = 1; w < nthreads; ++w) pool.emplace_back(guarded, w);
  guarded(0);
  for (auto& t : pool) t.join();
  if (failed) throw std::runtime_error(fail_msg);
  double squarings = 0, setup = 0;
  for (size_t w = 0; w < nthreads; ++w) {
    squarings += squarings_w[w];
    setup = std::max(setup, setup_w[w]);
  }
  // Tests cut off by the time limit.
  for (auto& j : jobs) {
    annie_batch_result& r = out[j->idx];
    if (j->m) {
      r.status = 2;
      r.iterations = j->m->iterations;
      std::snprintf(r.message, sizeof(r.message), "Test was aborted.");
    } else if (r.status == 0 && r.finished == 0 && r.message[0] == 0) {
      r.status = 2;
      std::snprintf(r.message, sizeof(r.message), "Test was not started.");
    }
  }
  if (st) {
    st->seconds = now();
    st->squarings = squarings;
    st->classes = int(classes.size());
    size_t slots = 0;
    for (auto& kv : classes) slots += kv.second.slots;
    st->slots = slots;
    st->setup_seconds = setup;
  }
  return 0;
}

}  // namespace annie_host

extern "C" int annie_batch_run(const annie_batch_test* tests, size_t count, const annie_batch_opts* opts,
                               annie_batch_result* results, annie_batch_stats* stats) {
  if ((!tests && count) || !opts || (!results && count)) return -1;
  try {
    return annie_host::batch_run(tests, count, *opts, results, stats);
  } catch (const std::exception& ex) {
    for (size_t i = 0; i < count; ++i)
      if (results[i].status == 0 && results[i].message[0] == 0) {
        results[i].status = -1;
        std::snprintf(results[i].message, sizeof(results[i].message), "Error: %s.", ex.what());
      }
    return -1;
  }
}
