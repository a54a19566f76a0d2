// ORACLE -- test infrastructure only.  Built by oracle/Makefile into
// oracle/_ref/libannie_ref.so from the reference's own header-only sources
// where they lie (-I <reference>/src; nothing is copied), and used by tests/
// and by bench.py's cpu_baseline leg as the reference CPU path.
//
// C ABI over the reference's transformGMP (src/transformGMP.h), transformCPU
// (src/transformCPU.h, K = 5 RNS NTT) and its test protocol annie::check
// (src/annie.h:217-430, built with CPU_TRANSFORM so that no OpenCL is needed).
#include <sys/stat.h>
#include <unistd.h>

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "annie.h"

namespace {

void from_limbs(mpz_t z, const uint64_t* l, size_t n) { mpz_import(z, n, -1, 8, 0, 0, l); }

long to_limbs(const mpz_t z, uint64_t* out, size_t cap) {
  const size_t need = mpz_sgn(z) == 0 ? 0 : (mpz_sizeinbase(z, 2) + 63) / 64;
  if (need > cap) return -1;
  size_t cnt = 0;
  if (need) mpz_export(out, &cnt, -1, 8, 0, 0, z);
  return long(need);
}

}  // namespace

extern "C" {

// kind 0: transformGMP, kind 1: transformCPU.
void* ref_create(int kind, const uint64_t* Nl, size_t nl, int32_t a) {
  mpz_t N;
  mpz_init(N);
  from_limbs(N, Nl, nl);
  transform* t = nullptr;
  if (kind == 0) t = new transformGMP(N, a);
  else t = new transformCPU(N, a);
  mpz_clear(N);
  return t;
}

int ref_init_failed(void* h) { return static_cast<transform*>(h)->init_failed() ? 1 : 0; }
void ref_set1(void* h) { static_cast<transform*>(h)->set1(); }
void ref_copy(void* h, int d, int s) { static_cast<transform*>(h)->copy(size_t(d), size_t(s)); }
void ref_square_mul(void* h, int32_t a) { static_cast<transform*>(h)->square_mul(a); }
void ref_mul(void* h, int d, int s) { static_cast<transform*>(h)->mul(size_t(d), size_t(s)); }

long ref_get(void* h, int s, uint64_t* out, size_t cap) {
  mpz_t z;
  mpz_init(z);
  static_cast<transform*>(h)->get(z, size_t(s));
  const long r = to_limbs(z, out, cap);
  mpz_clear(z);
  return r;
}

int ref_info(void* h, char* buf, size_t cap) {
  const std::string s = static_cast<transform*>(h)->get_info_short();
  if (cap == 0) return -1;
  std::strncpy(buf, s.c_str(), cap - 1);
  buf[cap - 1] = 0;
  return int(s.size());
}

void ref_destroy(void* h) { delete static_cast<transform*>(h); }

// annie::check on N, run in directory `workdir` (the protocol writes
// annie.log, results.txt and checkpoint files in the current directory).
// Returns the EReturn value (0 Success, 1 Failed, 2 Aborted).
int ref_check(const uint64_t* Nl, size_t nl, int is_euler, int32_t a, const char* workdir) {
  char cwd[4096];
  if (!getcwd(cwd, sizeof(cwd))) return -1;
  if (chdir(workdir) != 0) return -1;
  mpz_t N;
  mpz_init(N);
  from_limbs(N, Nl, nl);
  annie* an = new annie();
  const annie::EReturn r = an->check(N, "N", is_euler != 0, a, 0, true);
  delete an;
  mpz_clear(N);
  if (chdir(cwd) != 0) return -1;
  return int(r);
}

}  // extern "C"
