// Fixture generator (test infrastructure, run once in the build container):
// the Euler (Proth) test residue of N = k 2^n + 1 computed with GMP, for the
// golden file of the k*2^3,000,000+1 configuration.
//
// The reference's protocol (src/annie.h:249-253, 319-366): base a = smallest
// a >= 2 with Jacobi (a/N) = -1, perfect squares skipped; e = (N - 1) / 2 =
// k 2^(n-1); is_prp <=> a^e = -1 mod N; res64 = low 64 bits of a^(N-1) mod N
// (one squaring after the loop).  Here a^e = (a^k)^(2^(n-1)): b = a^k mod N by
// mpz_powm, then n - 1 modular squarings, then one more for res64.
//
// Each squaring is exact integer arithmetic with GMP:
//   * x^2 as a three-way split x = x2 2^(2s) + x1 2^s + x0 whose six partial
//     products run on six threads (plain mpz_mul each), summed with shifts;
//   * reduction mod N by the special form: x^2 = H 2^n + L (L < 2^n),
//     H = q k + r (0 <= r < k); since k 2^n = -1 mod N,
//     x^2 = L + r 2^n - q mod N, a value in (-N, N), so one conditional add of
//     N makes it canonical.
// Both steps are checked against plain mpz_mul + mpz_mod on the first
// iterations (and on random inputs at start-up); a mismatch aborts.
//
// Usage: proth_res64 k n state_file [report_iteration ...]
// Prints one line "iter <j> low64 <hex>" for each report iteration j (the
// canonical value of b^(2^j) mod N), then "a", "is_prp" and "res64" lines.
// The state file (iteration + current value) is rewritten every few minutes,
// and a run restarted with the same arguments resumes from it.
#include <gmp.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <thread>
#include <vector>

namespace {

struct Barrier {
  std::atomic<int> count{0};
  std::atomic<int> gen{0};
  int n;
  explicit Barrier(int nn) : n(nn) {}
  void wait() {
    const int g = gen.load(std::memory_order_acquire);
    if (count.fetch_add(1, std::memory_order_acq_rel) + 1 == n) {
      count.store(0, std::memory_order_relaxed);
      gen.fetch_add(1, std::memory_order_acq_rel);
    } else {
      while (gen.load(std::memory_order_acquire) == g) std::this_thread::yield();
    }
  }
};

// x^2 on six threads.  Thread t computes product t of the split.
struct Squarer {
  static constexpr int T = 6;
  mpz_t part[3], prod[T], *in = nullptr;
  mp_bitcnt_t s = 0;
  Barrier start{T + 1}, done{T + 1};
  std::atomic<bool> quit{false};
  std::vector<std::thread> th;

  Squarer() {
    for (auto& p : part) mpz_init(p);
    for (auto& p : prod) mpz_init(p);
    for (int t = 0; t < T; ++t) th.emplace_back([this, t] { run(t); });
  }
  ~Squarer() {
    quit = true;
    start.wait();
    for (auto& x : th) x.join();
  }
  void run(int t) {
    static const int A[T] = {0, 1, 2, 0, 0, 1}, B[T] = {0, 1, 2, 1, 2, 2};
    for (;;) {
      start.wait();
      if (quit) return;
      mpz_mul(prod[t], part[A[t]], part[B[t]]);
      done.wait();
    }
  }
  void square(mpz_t r, const mpz_t x) {
    const size_t bits = mpz_sizeinbase(x, 2);
    s = ((bits + 2) / 3 + 63) / 64 * 64;
    mpz_tdiv_r_2exp(part[0], x, s);
    mpz_tdiv_q_2exp(part[1], x, s);
    mpz_tdiv_q_2exp(part[2], part[1], s);
    mpz_tdiv_r_2exp(part[1], part[1], s);
    start.wait();
    done.wait();
    // x^2 = p00 + 2 p01 2^s + (p11 + 2 p02) 2^2s + 2 p12 2^3s + p22 2^4s
    mpz_set(r, prod[2]);                  // p22
    mpz_mul_2exp(r, r, s);
    mpz_addmul_ui(r, prod[5], 2);         // + 2 p12
    mpz_mul_2exp(r, r, s);
    mpz_add(r, r, prod[1]);               // + p11
    mpz_addmul_ui(r, prod[4], 2);         // + 2 p02
    mpz_mul_2exp(r, r, s);
    mpz_addmul_ui(r, prod[3], 2);         // + 2 p01
    mpz_mul_2exp(r, r, s);
    mpz_add(r, r, prod[0]);               // + p00
  }
};

struct Reducer {
  unsigned long k;
  mp_bitcnt_t n;
  mpz_t N, H, L, q;
  Reducer(unsigned long kk, mp_bitcnt_t nn, const mpz_t NN) : k(kk), n(nn) {
    mpz_inits(N, H, L, q, nullptr);
    mpz_set(N, NN);
  }
  ~Reducer() { mpz_clears(N, H, L, q, nullptr); }
  // r = x mod N for 0 <= x < N^2.
  void reduce(mpz_t r, const mpz_t x) {
    mpz_tdiv_r_2exp(L, x, n);
    mpz_tdiv_q_2exp(H, x, n);
    const unsigned long rem = mpz_tdiv_q_ui(q, H, k);
    mpz_set_ui(r, rem);
    mpz_mul_2exp(r, r, n);
    mpz_add(r, r, L);
    mpz_sub(r, r, q);
    if (mpz_sgn(r) < 0) mpz_add(r, r, N);
  }
};

void die(const char* m) {
  std::fprintf(stderr, "proth_res64: %s\n", m);
  std::exit(1);
}

bool load_state(const char* path, long& j, mpz_t x) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  long jj = 0;
  const bool ok = std::fread(&jj, sizeof(jj), 1, f) == 1 && mpz_inp_raw(x, f) != 0;
  std::fclose(f);
  if (ok) j = jj;
  return ok;
}

void save_state(const char* path, long j, const mpz_t x) {
  std::string tmp = std::string(path) + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) die("cannot write the state file");
  std::fwrite(&j, sizeof(j), 1, f);
  mpz_out_raw(f, x);
  std::fclose(f);
  std::rename(tmp.c_str(), path);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 4) die("usage: proth_res64 k n state_file [report_iteration ...]");
  const unsigned long k = std::strtoul(argv[1], nullptr, 10);
  const long n = std::strtol(argv[2], nullptr, 10);
  const char* state = argv[3];
  std::set<long> report;
  for (int i = 4; i < argc; ++i) report.insert(std::strtol(argv[i], nullptr, 10));
  if (k % 2 == 0 || n < 2) die("k must be odd, n >= 2");

  mpz_t N, b, x, y, t;
  mpz_inits(N, b, x, y, t, nullptr);
  mpz_set_ui(N, k);
  mpz_mul_2exp(N, N, mp_bitcnt_t(n));
  mpz_add_ui(N, N, 1);

  // Base of the Euler test (src/annie.h:250-253).
  long a = 2;
  while (mpz_si_kronecker(a, N) != -1) {
    ++a;
    long s = 0;
    while ((s + 1) * (s + 1) <= a) ++s;
    if (s * s == a) ++a;
  }
  std::printf("a %ld\n", a);
  std::fflush(stdout);

  Squarer sq;
  Reducer rd(k, mp_bitcnt_t(n), N);

  // Start-up self-check of the split squaring and the special reduction.
  {
    gmp_randstate_t rs;
    gmp_randinit_mt(rs);
    gmp_randseed_ui(rs, 12345);
    for (int i = 0; i < 8; ++i) {
      mpz_urandomm(x, rs, N);
      if (i == 0) mpz_sub_ui(x, N, 1);
      sq.square(y, x);
      mpz_mul(t, x, x);
      if (mpz_cmp(y, t) != 0) die("split squaring mismatch");
      rd.reduce(y, t);
      mpz_mod(t, t, N);
      if (mpz_cmp(y, t) != 0) die("special reduction mismatch");
    }
    gmp_randclear(rs);
  }

  // x = b^(2^j), starting from b = a^k mod N at j = 0.
  long j = 0;
  if (!load_state(state, j, x)) {
    mpz_set_ui(b, k);
    mpz_set_ui(t, unsigned(a));
    mpz_powm(x, t, b, N);
    j = 0;
    save_state(state, j, x);
  } else {
    std::fprintf(stderr, "resuming at iteration %ld\n", j);
  }
  auto last = std::chrono::steady_clock::now();
  const auto t0 = last;
  const long total = n;  // n - 1 squarings to a^e, one more to a^(N-1)
  bool is_prp = false;
  for (; j < total;) {
    if (j == n - 1) {
      mpz_sub_ui(t, N, 1);
      is_prp = mpz_cmp(x, t) == 0;
      std::printf("is_prp %d\n", is_prp ? 1 : 0);
      std::fflush(stdout);
    }
    sq.square(y, x);
    if (j < 64) {
      mpz_mul(t, x, x);
      if (mpz_cmp(y, t) != 0) die("split squaring mismatch");
      mpz_mod(t, t, N);
      rd.reduce(x, y);
      if (mpz_cmp(x, t) != 0) die("special reduction mismatch");
    } else {
      rd.reduce(x, y);
    }
    ++j;
    if (report.count(j)) {
      std::printf("iter %ld low64 %016llx\n", j, (unsigned long long)(mpz_size(x) ? mpz_getlimbn(x, 0) : 0));
      std::fflush(stdout);
    }
    const auto now = std::chrono::steady_clock::now();
    if (std::chrono::duration<double>(now - last).count() > 300) {
      save_state(state, j, x);
      last = now;
      const double el = std::chrono::duration<double>(now - t0).count();
      std::fprintf(stderr, "iteration %ld of %ld, %.1f s\n", j, total, el);
    }
  }
  std::printf("res64 %016llx\n", (unsigned long long)(mpz_size(x) ? mpz_getlimbn(x, 0) : 0));
  std::printf("done\n");
  mpz_clears(N, b, x, y, t, nullptr);
  return 0;
}
