// `annie` command-line tool: probable-primality test of an expression on the
// gfx950 engine (restates src/main.cpp without PARI/GP).
//
//   annie -q "<expression>" [-fermat <a> | -euler] [-d <device>] [-w <dir>]
//   annie -f <file> ...      (first line of the file is the expression)
//   annie -h [-d <device>]   (Gerbicz-Li self-check and ms/bit at 3*2^e+1 sizes)
//
// Expressions: integers, + - * / % ^, parentheses, postfix n! (factorial)
// and n# (primorial), fibonacci(n), lucas(n), sqrtnint(x, k), nextprime(x).
#include <gmp.h>

#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "annie_hip.h"
#include "annie_host.h"

namespace {

class Parser {
  const std::string s_;
  size_t i_ = 0;

  void ws() {
    while (i_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[i_]))) ++i_;
  }
  bool eat(char c) {
    ws();
    if (i_ < s_.size() && s_[i_] == c) {
      ++i_;
      return true;
    }
    return false;
  }
  [[noreturn]] void error(const std::string& m) const {
    throw std::runtime_error("expression: " + m + " at position " + std::to_string(i_));
  }
  static unsigned long small(const mpz_t x) {
    if (mpz_sgn(x) < 0 || !mpz_fits_ulong_p(x)) throw std::runtime_error("expression: argument too large");
    return mpz_get_ui(x);
  }

  void expr(mpz_t r) {
    term(r);
    for (;;) {
      mpz_t t;
      if (eat('+')) {
        mpz_init(t);
        term(t);
        mpz_add(r, r, t);
      } else if (eat('-')) {
        mpz_init(t);
        term(t);
        mpz_sub(r, r, t);
      } else {
        return;
      }
      mpz_clear(t);
    }
  }
  // Precedence as in PARI/GP (the reference's evaluator, src/main.cpp:252):
  // unary minus binds looser than '^' (-2^2 = -4) and tighter than '*'
  // (2*-3 = -6); '^' is right associative and takes a signed exponent.
  void term(mpz_t r) {
    unary(r);
    for (;;) {
      mpz_t t;
      char op;
      if (eat('*')) op = '*';
      else if (eat('/')) op = '/';
      else if (eat('%')) op = '%';
      else return;
      mpz_init(t);
      unary(t);
      if (op == '*') mpz_mul(r, r, t);
      else if (mpz_sgn(t) == 0) error("division by zero");
      else if (op == '/') mpz_tdiv_q(r, r, t);
      else mpz_mod(r, r, t);
      mpz_clear(t);
    }
  }
  void unary(mpz_t r) {
    if (eat('-')) {
      unary(r);
      mpz_neg(r, r);
      return;
    }
    if (eat('+')) {
      unary(r);
      return;
    }
    power(r);
  }
  void power(mpz_t r) {
    postfix(r);
    if (eat('^')) {
      mpz_t e;
      mpz_init(e);
      unary(e);  // right associative: 2^3^2 = 2^9
      if (mpz_sgn(e) < 0) error("negative exponent (not an integer)");
      mpz_pow_ui(r, r, small(e));
      mpz_clear(e);
    }
  }
  void postfix(mpz_t r) {
    primary(r);
    for (;;) {
      if (eat('!')) mpz_fac_ui(r, small(r));
      else if (eat('#')) mpz_primorial_ui(r, small(r));
      else return;
    }
  }
  void primary(mpz_t r) {
    ws();
    if (eat('(')) {
      expr(r);
      if (!eat(')')) error("')' expected");
      return;
    }
    if (i_ < s_.size() && std::isdigit(static_cast<unsigned char>(s_[i_]))) {
      size_t j = i_;
      while (j < s_.size() && std::isdigit(static_cast<unsigned char>(s_[j]))) ++j;
      mpz_set_str(r, s_.substr(i_, j - i_).c_str(), 10);
      i_ = j;
      return;
    }
    size_t j = i_;
    while (j < s_.size() && std::isalpha(static_cast<unsigned char>(s_[j]))) ++j;
    const std::string f = s_.substr(i_, j - i_);
    if (f.empty()) error("unexpected character");
    i_ = j;
    if (!eat('(')) error("'(' expected after " + f);
    std::vector<mpz_t*> args;
    do {
      mpz_t* a = new mpz_t[1];
      mpz_init(*a);
      expr(*a);
      args.push_back(a);
    } while (eat(','));
    if (!eat(')')) error("')' expected");
    auto need = [&](size_t k) {
      if (args.size() != k) error(f + ": wrong number of arguments");
    };
    if (f == "fibonacci") {
      need(1);
      mpz_fib_ui(r, small(*args[0]));
    } else if (f == "lucas") {
      need(1);
      mpz_lucnum_ui(r, small(*args[0]));
    } else if (f == "sqrtnint") {
      need(2);
      mpz_root(r, *args[0], small(*args[1]));
    } else if (f == "nextprime") {
      need(1);
      mpz_sub_ui(r, *args[0], 1);
      mpz_nextprime(r, r);
    } else {
      error("unknown function " + f);
    }
    for (mpz_t* a : args) {
      mpz_clear(*a);
      delete[] a;
    }
  }

 public:
  explicit Parser(std::string s) : s_(std::move(s)) {}
  void eval(mpz_t r) {
    expr(r);
    ws();
    if (i_ != s_.size()) error("trailing characters");
  }
};

std::vector<uint64_t> limbs(const mpz_t z) {
  std::vector<uint64_t> l((mpz_sizeinbase(z, 2) + 63) / 64 + 1, 0);
  size_t cnt = 0;
  mpz_export(l.data(), &cnt, -1, 8, 0, 0, z);
  l.resize(cnt ? cnt : 1);
  return l;
}

void on_signal(int) { annie_request_quit(); }

// Bench mode (src/annie.h:432-555): Gerbicz-Li self-check on a fixed exponent
// and ms per squaring for N = 3*2^e + 1.
int bench(int device) {
  static const unsigned long sizes[] = {33216, 83048, 166092, 332190, 664381, 1162672, 1660962,
                                        3321924, 6643854, 16609638, 33219278, 66438557, 159200000};
  mpz_t n;
  mpz_init(n);
  for (unsigned long e : sizes) {
    mpz_set_ui(n, 3);
    mpz_mul_2exp(n, n, e);
    mpz_add_ui(n, n, 1);
    const std::vector<uint64_t> l = limbs(n);
    annie_hip_t* h = nullptr;
    if (annie_hip_create(l.data(), l.size(), 3, device, 10, &h) != ANNIE_OK) {
      std::printf("3*2^%lu + 1: %s\n", e, annie_hip_last_error());
      annie_hip_destroy(h);
      continue;
    }
    int K, pck, LN, lbase;
    annie_hip_info(h, &K, &pck, &LN, &lbase);
    std::vector<uint8_t> bits(2000);
    for (size_t i = 0; i < bits.size(); ++i) bits[i] = (i & 1) ? 1 : 0;
    annie_hip_set1(h);
    annie_hip_square_mul_bits(h, bits.data(), 50, 3);
    float ms = 0;
    annie_hip_time_square_mul_bits(h, bits.data(), bits.size(), 3, &ms);
    const double per = ms / double(bits.size());
    std::printf("3*2^%lu + 1, %d-prime NTT of length %zu, packed-base = 2^{%dx%d}, %.4f ms/bit, full test ~ %.1f s\n",
                e, K, size_t(1) << LN, pck, lbase, per, per * double(e) * 1e-3);
    std::fflush(stdout);
    annie_hip_destroy(h);
    if (std::getenv("ANNIE_BENCH_MAX_BITS") && e >= std::strtoul(std::getenv("ANNIE_BENCH_MAX_BITS"), nullptr, 10))
      break;
  }
  mpz_clear(n);
  return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
  std::signal(SIGINT, on_signal);
  std::signal(SIGTERM, on_signal);
  std::string expression, filename, workdir = ".";
  int device = 0, a = 0;
  bool euler = false, hbench = false;
  try {
    for (int i = 1; i < argc; ++i) {
      const std::string arg = argv[i];
      auto val = [&](const std::string& flag) -> std::string {
        if (arg == flag && i + 1 < argc) return argv[++i];
        return arg.substr(flag.size());
      };
      if (arg.rfind("-q", 0) == 0) expression = val("-q");
      else if (arg.rfind("-fermat", 0) == 0) a = std::atoi(val("-fermat").c_str());
      else if (arg.rfind("-euler", 0) == 0) euler = true;
      else if (arg.rfind("-f", 0) == 0) filename = val("-f");
      else if (arg.rfind("-d", 0) == 0) device = std::atoi(val("-d").c_str());
      else if (arg.rfind("-w", 0) == 0) workdir = val("-w");
      else if (arg.rfind("-h", 0) == 0) hbench = true;
    }
    if (euler && a != 0) throw std::runtime_error("-euler used with an incompatible option (-fermat)");
    if (hbench) return bench(device);
    if (!filename.empty()) {
      std::ifstream f(filename);
      if (!f.is_open()) throw std::runtime_error("cannot open input file");
      std::getline(f, expression);
    }
    if (expression.empty()) {
      std::printf("Usage: annie -q <expression> [-fermat <a> | -euler] [-d <device>] [-w <dir>]\n"
                  "       annie -f <file> ...\n       annie -h [-d <device>]\n");
      return 0;
    }
    mpz_t n;
    mpz_init(n);
    Parser(expression).eval(n);
    if (mpz_sgn(n) <= 0) throw std::runtime_error("the expression must be a positive integer");
    const std::vector<uint64_t> l = limbs(n);
    mpz_clear(n);
    annie_check_opts o{};
    o.is_euler = euler;
    o.a_ref = a;
    o.device = device;
    o.gmp_threshold_digits = 10000;
    o.min_ln = 10;
    o.workdir = workdir.c_str();
    o.checkpoint_seconds = 600;
    o.quit_after = -1;
    o.corrupt_at = -1;
    o.verbose = 1;
    o.expression = expression.c_str();
    annie_check_result r{};
    annie_check(l.data(), l.size(), &o, &r);
    if (r.status == 1) std::printf("Validation failed!\n");
    else if (r.status == 2) std::printf("Test was aborted.\n");
    else if (r.status < 0) throw std::runtime_error(r.message);
    return r.status == 1 ? 1 : 0;
  } catch (const std::exception& ex) {
    std::fprintf(stderr, "\nError: %s.\n", ex.what());
    return 1;
  }
}
