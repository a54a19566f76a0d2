// ORACLE -- test infrastructure only.  Built by oracle/Makefile into
// oracle/_ref/ref_plan from the reference's own sources where they lie
// (-I <reference>/src; nothing is copied).  Runs the reference's size
// selection create_transformGPU (src/transformGPU.h:822-885) on N (hex, read
// from standard input) and base a (argument): the function prints its candidate line per K and
// the chosen K_min (src/transformGPU.h:838-843) before it constructs the
// OpenCL engine.  That construction needs a GPU and an OpenCL device this
// container does not have, so the process may end abnormally after the
// selection lines; the caller reads stdout only (every line is flushed).
#include <gmp.h>

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <iterator>
#include <string>

#include "transformGPU.h"

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: ref_plan <a> < N_in_hex\n");
    return 2;
  }
  std::string hex((std::istreambuf_iterator<char>(std::cin)), std::istreambuf_iterator<char>());
  while (!hex.empty() && (hex.back() == '\n' || hex.back() == ' ')) hex.pop_back();
  mpz_t N;
  mpz_init(N);
  if (mpz_set_str(N, hex.c_str(), 16) != 0) return 2;
  const int32_t a = int32_t(std::atoi(argv[1]));
  std::cout << std::unitbuf;
  try {
    transform* t = create_transformGPU(N, a, 0);
    delete t;
  } catch (const std::exception& e) {
    std::cout << "exception: " << e.what() << std::endl;
  }
  mpz_clear(N);
  return 0;
}
