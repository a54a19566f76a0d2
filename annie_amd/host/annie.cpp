// Test protocol host: Fermat / Euler probable-primality test with the
// Gerbicz-Li exponentiation proof check, checkpoints and resume, driving a
// `transform` (the gfx950 engine, or GMP for small N).
//
// Restates annie::check (src/annie.h:217-430) and the checkpoint file format
// of src/file.h (running CRC-32 with the trailer ~crc ^ 0xa23777ac) and
// src/annie.h:125-197 (version, i, elapsed, transform payload).
#include <gmp.h>
#include <sys/stat.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iomanip>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "annie_host.h"
#include "transform.h"

namespace annie_host {

static std::atomic<bool> g_quit(false);

// CRC-32 (reflected polynomial 0xedb88320), as src/file.h:56-90.
static uint32_t crc32_update(uint32_t crc, const void* buf, size_t len) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t r = i;
      for (int j = 0; j < 8; ++j) r = (r & 1) ? (r >> 1) ^ 0xedb88320u : (r >> 1);
      table[i] = r;
    }
    init = true;
  }
  uint32_t c = ~crc;
  const uint8_t* p = static_cast<const uint8_t*>(buf);
  for (size_t i = 0; i < len; ++i) c = (c >> 8) ^ table[(c & 0xff) ^ p[i]];
  return ~c;
}

// Exponent block of the Gerbicz-Li proof (src/annie.h:95-101).
static int gl_block(size_t size) {
  int lg = -1;
  for (uint32_t v = uint32_t(size); v; v >>= 1) ++lg;
  const size_t L = size_t(2) << (lg / 2);
  return int((size - 1) / L) + 1;
}

static size_t decimal_digits(const mpz_t n) {
  std::string s(mpz_sizeinbase(n, 10) + 2, '\0');
  mpz_get_str(&s[0], 10, n);
  return std::strlen(s.c_str());
}

static std::string hex64(uint64_t u) {
  std::ostringstream ss;
  ss << std::uppercase << std::hex << std::setfill('0') << std::setw(16) << u;
  return ss.str();
}

static std::string checkpoint_name(const mpz_t n, size_t nsize) {
  long e = 0;
  const double d = mpz_get_d_2exp(&e, n);
  const uint64_t hash = uint64_t(d * double(uint64_t(1) << 52)) * nsize + mpz_getlimbn(n, 0);
  return "a_" + hex64(hash) + ".chk";
}

// Checkpoint file: int32 version = 1, int32 i, double elapsed, payload, CRC.
static bool write_checkpoint(const std::string& path, int i, double elapsed, const std::vector<char>& payload) {
  const std::string tmp = path + ".new", old = path + ".old";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return false;
  uint32_t crc = 0;
  auto put = [&](const void* p, size_t n) {
    crc = crc32_update(crc, p, n);
    return std::fwrite(p, 1, n, f) == n;
  };
  const int32_t version = 1, ii = i;
  bool ok = put(&version, 4) && put(&ii, 4) && put(&elapsed, 8) && put(payload.data(), payload.size());
  const uint32_t trailer = ~crc ^ 0xa23777acu;
  ok = ok && std::fwrite(&trailer, 1, 4, f) == 4;
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return false;
  std::remove(old.c_str());
  struct stat st;
  if (stat(path.c_str(), &st) == 0) std::rename(path.c_str(), old.c_str());
  return std::rename(tmp.c_str(), path.c_str()) == 0;
}

// Returns 0 on success, -1 if absent, -2 if invalid.
static int read_checkpoint_file(const std::string& path, size_t payload_size, int& i, double& elapsed,
                                std::vector<char>& payload) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return -1;
  uint32_t crc = 0;
  auto get = [&](void* p, size_t n) {
    const bool r = std::fread(p, 1, n, f) == n;
    crc = crc32_update(crc, p, n);
    return r;
  };
  int32_t version = 0, ii = 0;
  payload.assign(payload_size, 0);
  bool ok = get(&version, 4) && version == 1 && get(&ii, 4) && get(&elapsed, 8) && get(payload.data(), payload_size);
  uint32_t trailer = 0;
  ok = ok && std::fread(&trailer, 1, 4, f) == 4 && trailer == (~crc ^ 0xa23777acu);
  std::fclose(f);
  if (!ok) return -2;
  i = ii;
  return 0;
}

static void append_file(const std::string& path, const std::string& text) {
  FILE* f = std::fopen(path.c_str(), "a");
  if (!f) return;
  std::fputs(text.c_str(), f);
  std::fclose(f);
}

struct Run {
  const annie_check_opts& o;
  annie_check_result& r;
  std::string dir;
  void log(const std::string& s) const {
    if (o.verbose) std::fputs(s.c_str(), stdout), std::fflush(stdout);
    if (!dir.empty()) append_file(dir + "/annie.log", s);
  }
  void result(const std::string& s) const {
    std::snprintf(r.message, sizeof(r.message), "%s", s.c_str());
    if (!dir.empty()) append_file(dir + "/results.txt", s + "\n");
    log(s + "\n");
  }
};

static int check(const mpz_t n, const annie_check_opts& o, annie_check_result& r) {
  using clock = std::chrono::steady_clock;
  const auto t_start = clock::now();
  Run run{o, r, o.workdir ? std::string(o.workdir) : std::string()};
  const size_t nsize = decimal_digits(n);
  std::string ns;
  {
    std::string s(mpz_sizeinbase(n, 10) + 2, '\0');
    mpz_get_str(&s[0], 10, n);
    s = s.c_str();
    ns = (s.size() <= 40) ? s : s.substr(0, 8) + "[" + std::to_string(s.size() - 16) + " other digits]" +
                                     s.substr(s.size() - 8);
    if (o.expression && std::strlen(o.expression) <= 40) ns = o.expression;
  }

  if (nsize <= 15) {  // src/annie.h:222-237
    const int pp = mpz_probab_prime_p(n, 20);
    r.status = 4;
    r.is_prp = pp != 0;
    run.result(ns + (pp == 0 ? " is composite." : (pp == 1 ? " is a probable prime." : " is prime.")));
    return 0;
  }
  if (mpz_even_p(n)) {  // src/annie.h:239-247
    r.status = 4;
    run.result(ns + " is even.");
    return 0;
  }

  int32_t a = 2;  // src/annie.h:249-254
  if (o.is_euler) {
    while (mpz_si_kronecker(a, n) != -1) {
      ++a;
      const int32_t s = int32_t(std::lround(std::sqrt(double(a))));
      if (s * s == a) ++a;
    }
  } else if (o.a_ref != 0) {
    a = o.a_ref;
  }
  r.a = a;
  {
    std::ostringstream ss;
    ss << "Running " << (o.is_euler ? "Euler" : "Fermat") << " probable primality test of " << ns << " to base " << a
       << " [" << nsize << " digits]\n";
    run.log(ss.str());
  }

  const int thr = o.gmp_threshold_digits > 0 ? o.gmp_threshold_digits : 10000;
  std::unique_ptr<transform> T(nsize < size_t(thr) ? create_transform_gmp(n, a)
                                                    : create_transform_hip(n, a, o.device, o.min_ln > 0 ? o.min_ln : 10));
  if (T->init_failed()) {  // src/annie.h:269-283
    mpz_t f;
    mpz_init(f);
    T->get_factor(f);
    std::string s(mpz_sizeinbase(f, 10) + 2, '\0');
    mpz_get_str(&s[0], 10, f);
    mpz_clear(f);
    r.status = 3;
    run.result(std::string(s.c_str()) + " divides " + ns + ".");
    return 0;
  }
  run.log(T->get_info() + "\n");

  const std::string cpath = run.dir.empty() ? std::string() : run.dir + "/" + checkpoint_name(n, nsize);
  mpz_t e;
  mpz_init(e);
  mpz_sub_ui(e, n, 1);
  if (o.is_euler) mpz_tdiv_q_2exp(e, e, 1);
  const int B = gl_block(mpz_sizeinbase(e, 2));

  int i_start = int(mpz_sizeinbase(e, 2)) - 1;
  double restored = 0;
  bool found = false;
  if (!cpath.empty() && T->has_checkpoint()) {  // src/annie.h:140-160: current, then .old
    std::vector<char> payload = T->save_checkpoint();  // size probe
    const size_t psize = payload.size();
    for (const std::string& p : {cpath, cpath + ".old"}) {
      int ci = 0;
      double el = 0;
      const int rc = read_checkpoint_file(p, psize, ci, el, payload);
      if (rc == -2) run.log("Error: " + p + ": invalid checkpoint.\n");
      if (rc == 0 && T->read_checkpoint(payload)) {
        i_start = ci;
        restored = el;
        found = true;
        break;
      }
    }
  }
  if (found) {
    run.log("Resuming from a checkpoint.\n");
    r.resumed = 1;
  } else {
    T->set1();
    T->copy(transform::Reg1, transform::Reg0);  // d(t)
  }

  auto elapsed = [&]() { return restored + std::chrono::duration<double>(clock::now() - t_start).count(); };
  auto last_ckpt = clock::now();
  // Progress line every 10 s of display time (src/annie.h:103-120, 328-331):
  // fraction done, remaining / expected time, ms per loop iteration.
  const int i_top = int(mpz_sizeinbase(e, 2)) - 1;
  auto last_disp = clock::now();
  int disp_i = i_start;
  auto progress = [&](int i) {
    const double dt = std::chrono::duration<double>(clock::now() - last_disp).count();
    if (dt < 10.0 || disp_i == i) return;
    const double frac = double(i_top - i) / double(std::max(i_top, 1));
    const double per = dt / double(disp_i - i);
    std::printf("%.3g%% done, %.0f s / %.0f s remaining, %.4g ms/bit.\n", frac * 100.0, per * i,
                frac > 0 ? elapsed() / frac : 0.0, per * 1e3);
    std::fflush(stdout);
    disp_i = i;
    last_disp = clock::now();
  };
  long iters = 0;
  for (int i = i_start; i >= 0; --i) {  // hot loop, src/annie.h:319-341
    if (o.verbose && (i & 1023) == 0) progress(i);
    if (g_quit.load(std::memory_order_relaxed) || (o.quit_after >= 0 && iters == o.quit_after)) {
      if (!cpath.empty() && T->has_checkpoint()) write_checkpoint(cpath, i, elapsed(), T->save_checkpoint());
      r.status = 2;
      r.iterations = iters;
      r.seconds = elapsed();
      std::snprintf(r.message, sizeof(r.message), "Test was aborted.");
      mpz_clear(e);
      return 0;
    }
    if (!cpath.empty() && T->has_checkpoint() && o.checkpoint_seconds > 0 && (i & 1023) == 0 &&
        std::chrono::duration<double>(clock::now() - last_ckpt).count() > o.checkpoint_seconds) {
      write_checkpoint(cpath, i, elapsed(), T->save_checkpoint());
      last_ckpt = clock::now();
    }
    T->square_mul(mpz_tstbit(e, mp_bitcnt_t(i)) ? a : 1);
    if ((i % B == 0) && (i / B != 0)) T->mul(transform::Reg1, transform::Reg0);
    if (o.corrupt_at >= 0 && iters == o.corrupt_at) T->corrupt(transform::Reg0);
    ++iters;
  }

  T->copy(transform::Reg2, transform::Reg0);  // d(t+1) = d(t) * result
  T->mul(transform::Reg2, transform::Reg1);

  mpz_t t, res, r1, r2;
  mpz_inits(t, res, r1, r2, nullptr);
  T->get(t, transform::Reg0);
  if (o.is_euler) {  // result ?= -1, then a^(N-1) for the residue (src/annie.h:349-356)
    mpz_add_ui(t, t, 1);
    if (mpz_cmp(t, n) >= 0) mpz_sub(t, t, n);
    T->square_mul(1);
  } else {
    mpz_sub_ui(t, t, 1);
  }
  r.is_prp = mpz_sgn(t) == 0;
  T->get(t, transform::Reg0);
  r.res64 = mpz_size(t) ? uint64_t(mpz_getlimbn(t, 0)) : 0;

  mpz_set_ui(res, 0);  // res = sum of the B-bit chunks of e (src/annie.h:368-375)
  while (mpz_sgn(e) != 0) {
    mpz_tdiv_r_2exp(t, e, mp_bitcnt_t(B));
    mpz_add(res, res, t);
    mpz_tdiv_q_2exp(e, e, mp_bitcnt_t(B));
  }
  const int rs = int(mpz_sizeinbase(res, 2));
  T->set1();  // a^(res - 2^B) ... (src/annie.h:377-404)
  for (int i = rs - 1; i >= B; --i) T->square_mul(mpz_tstbit(res, mp_bitcnt_t(i)) ? a : 1);
  T->mul(transform::Reg0, transform::Reg1);
  for (int i = B - 1; i >= 0; --i) T->square_mul((i < rs && mpz_tstbit(res, mp_bitcnt_t(i))) ? a : 1);
  T->get(r1, transform::Reg0);
  T->get(r2, transform::Reg2);
  r.gl_ok = mpz_cmp(r1, r2) == 0;
  mpz_clears(t, res, r1, r2, e, nullptr);

  r.iterations = iters;
  r.seconds = elapsed();
  std::ostringstream ss;
  if (r.gl_ok) {
    if (r.is_prp)
      ss << ns << " is a " << (o.is_euler ? "Euler" : "Fermat") << " probable prime, a = " << a << ".";
    else
      ss << ns << " is composite, res64 = " << hex64(r.res64) << ", a = " << a << ".";
    r.status = 0;
  } else {
    ss << ns << ": validation failed!";
    r.status = 1;
  }
  run.result(ss.str());
  if (!cpath.empty()) {  // src/annie.h:428
    std::remove(cpath.c_str());
    std::remove((cpath + ".old").c_str());
  }
  return 0;
}

}  // namespace annie_host

extern "C" {

int annie_check(const uint64_t* Nl, size_t nl, const annie_check_opts* opts, annie_check_result* res) {
  if (!Nl || !opts || !res) return -1;
  std::memset(res, 0, sizeof(*res));
  mpz_t n;
  mpz_init(n);
  mpz_import(n, nl, -1, 8, 0, 0, Nl);
  int rc = 0;
  try {
    rc = annie_host::check(n, *opts, *res);
  } catch (const std::exception& ex) {
    res->status = -1;
    std::snprintf(res->message, sizeof(res->message), "Error: %s.", ex.what());
    rc = -1;
  }
  mpz_clear(n);
  return rc;
}

void annie_request_quit(void) { annie_host::g_quit.store(true); }
void annie_clear_quit(void) { annie_host::g_quit.store(false); }

int annie_checkpoint_name(const uint64_t* Nl, size_t nl, char* buf, size_t cap) {
  mpz_t n;
  mpz_init(n);
  mpz_import(n, nl, -1, 8, 0, 0, Nl);
  const std::string s = annie_host::checkpoint_name(n, annie_host::decimal_digits(n));
  mpz_clear(n);
  if (!buf || cap <= s.size()) return -1;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return 0;
}

uint32_t annie_crc32(uint32_t crc, const void* buf, size_t len) { return annie_host::crc32_update(crc, buf, len); }

}  // extern "C"
