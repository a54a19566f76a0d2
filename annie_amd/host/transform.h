// The register-machine interface the test protocol drives (the reference's
// `transform`, src/transform.h:18-45): three user registers holding values in
// Montgomery-McLaughlin-Gallot-Woltman form, squared / multiplied in place.
#pragma once
#include <gmp.h>

#include <cstdint>
#include <string>
#include <vector>

namespace annie_host {

class transform {
 public:
  static const size_t Reg0 = 0, Reg1 = 1, Reg2 = 2;

  virtual ~transform() {}
  virtual void set1() = 0;                              // r0 = 1
  virtual void get(mpz_t z, size_t src) = 0;            // z = r_src, in [0, N)
  virtual void copy(size_t dst, size_t src) = 0;        // r_dst = r_src
  virtual void square_mul(int32_t a) = 0;               // r0 = a r0^2
  virtual void square_mul_bits(const std::vector<uint8_t>& bits, int32_t a) {
    for (uint8_t b : bits) square_mul(b ? a : 1);
  }
  virtual void mul(size_t dst, size_t src) = 0;         // r_dst *= r_src
  virtual bool init_failed() const = 0;
  virtual void get_factor(mpz_t z) const = 0;
  virtual bool has_checkpoint() const = 0;
  virtual bool read_checkpoint(const std::vector<char>& payload) = 0;  // r0 and r1
  virtual std::vector<char> save_checkpoint() = 0;                      // r0 and r1
  virtual std::string get_info() const = 0;
  // Test hook: flips one word of a user register (failure-detection tests).
  virtual void corrupt(size_t reg) { (void)reg; }
};

// Exact MMGW with GMP, used below 10,000 digits like the reference
// (src/annie.h:31-35, 263; src/transformGMP.h).
transform* create_transform_gmp(const mpz_t N, int32_t a);

// The gfx950 engine through its C ABI (replaces create_transformGPU,
// src/transformGPU.h:822-885).  Throws std::runtime_error on failure.
transform* create_transform_hip(const mpz_t N, int32_t a, int device, int min_ln);

}  // namespace annie_host
