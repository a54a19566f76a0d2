// Implementations of the host `transform` interface: the GMP path for small N
// and the adapter over the gfx950 engine's C ABI.
#include <gmp.h>

#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "annie_hip.h"
#include "transform.h"

namespace annie_host {

namespace {

// MMGW with GMP (src/transformGMP.h): R = 2^nb - 1, Q = 2^nb + 1, 2^nb > 16 a N.
class transform_gmp : public transform {
  mp_bitcnt_t nb_ = 0;
  bool failed_ = false;
  mpz_t N_, Ninv_, Q_, one_, tq_, t_, reg_[3];

  // z mod R, folding the bits above nb; result in [0, R].
  void mod_R(mpz_t z) {
    while (mpz_sizeinbase(z, 2) > nb_) {
      mpz_tdiv_q_2exp(t_, z, nb_);
      mpz_tdiv_r_2exp(z, z, nb_);
      mpz_add(z, z, t_);
    }
  }
  // z mod Q for z >= 0: fold 2nb-bit chunks, then low - high, lifted to [0, Q).
  void mod_Q(mpz_t z) {
    while (mpz_sizeinbase(z, 2) > 2 * nb_) {
      mpz_tdiv_q_2exp(t_, z, 2 * nb_);
      mpz_tdiv_r_2exp(z, z, 2 * nb_);
      mpz_add(z, z, t_);
    }
    mpz_tdiv_q_2exp(t_, z, nb_);
    mpz_tdiv_r_2exp(z, z, nb_);
    mpz_sub(z, z, t_);
    if (mpz_sgn(z) < 0) mpz_add(z, z, Q_);
  }
  // x <- ((x N^-1 mod R) + 2) N - T_Q mod Q (doc/annie.tex:151-164).
  void reduce(mpz_t x) {
    mpz_mul(x, x, Ninv_);
    mod_R(x);
    mpz_add_ui(x, x, 2);
    mpz_mul(x, x, N_);
    mod_Q(x);
    mpz_sub(x, x, tq_);
    if (mpz_sgn(x) < 0) mpz_add(x, x, Q_);
  }

 public:
  transform_gmp(const mpz_t N, int32_t a) {
    mpz_inits(N_, Ninv_, Q_, one_, tq_, t_, reg_[0], reg_[1], reg_[2], nullptr);
    mpz_set(N_, N);
    int abits = 0;
    for (uint32_t v = uint32_t(a); v; v >>= 1) ++abits;
    nb_ = mpz_sizeinbase(N, 2) + size_t(abits) + 4;
    mpz_t R;
    mpz_init(R);
    mpz_set_ui(R, 1);
    mpz_mul_2exp(R, R, nb_);
    mpz_add_ui(Q_, R, 1);
    mpz_sub_ui(R, R, 1);
    if (mpz_invert(Ninv_, N, R) == 0) {
      mpz_gcd(Ninv_, N, R);
      failed_ = true;
    } else {
      mpz_mod(one_, R, N);
      if (mpz_odd_p(one_)) mpz_add(one_, one_, N);
      mpz_tdiv_q_2exp(one_, one_, 1);
    }
    mpz_clear(R);
  }
  ~transform_gmp() override { mpz_clears(N_, Ninv_, Q_, one_, tq_, t_, reg_[0], reg_[1], reg_[2], nullptr); }

  void set1() override { mpz_set(reg_[0], one_); }
  void get(mpz_t z, size_t src) override {
    mpz_set(tq_, reg_[src]);
    mpz_set(z, reg_[src]);
    reduce(z);
    mpz_mod(z, z, N_);
  }
  void copy(size_t dst, size_t src) override { mpz_set(reg_[dst], reg_[src]); }
  void square_mul(int32_t a) override {
    mpz_mul_si(tq_, reg_[0], a);
    mpz_mul(tq_, tq_, reg_[0]);
    mod_Q(tq_);
    mpz_mul(reg_[0], reg_[0], reg_[0]);
    mpz_mul_si(reg_[0], reg_[0], a);
    mod_R(reg_[0]);
    reduce(reg_[0]);
  }
  void mul(size_t dst, size_t src) override {
    mpz_mul(tq_, reg_[dst], reg_[src]);
    mod_Q(tq_);
    mpz_mul(reg_[dst], reg_[dst], reg_[src]);
    mod_R(reg_[dst]);
    reduce(reg_[dst]);
  }
  bool init_failed() const override { return failed_; }
  void get_factor(mpz_t z) const override { mpz_set(z, Ninv_); }
  bool has_checkpoint() const override { return false; }
  bool read_checkpoint(const std::vector<char>&) override { return false; }
  std::vector<char> save_checkpoint() override { return {}; }
  std::string get_info() const override {
    std::ostringstream ss;
    ss << "Montgomery-McLaughlin-Gallot-Woltman reduction of length " << nb_ << " bits";
    return ss.str();
  }
  void corrupt(size_t reg) override { mpz_add_ui(reg_[reg], reg_[reg], 1); }
};

std::vector<uint64_t> to_limbs(const mpz_t z) {
  std::vector<uint64_t> l((mpz_sizeinbase(z, 2) + 63) / 64 + 1, 0);
  size_t cnt = 0;
  mpz_export(l.data(), &cnt, -1, 8, 0, 0, z);
  l.resize(cnt ? cnt : 1);
  return l;
}

void from_limbs(mpz_t z, const std::vector<uint64_t>& l, size_t n) { mpz_import(z, n, -1, 8, 0, 0, l.data()); }

// The gfx950 engine behind the interface.
class transform_hip : public transform {
  annie_hip_t* h_ = nullptr;
  bool failed_ = false;
  size_t nlimbs_ = 0;
  int K_ = 0, pck_ = 0, LN_ = 0, lbase_ = 0;

  void check(int rc) const {
    if (rc != ANNIE_OK) throw std::runtime_error(annie_hip_last_error());
  }

 public:
  transform_hip(const mpz_t N, int32_t a, int device, int min_ln) {
    const std::vector<uint64_t> l = to_limbs(N);
    nlimbs_ = l.size() + 1;
    const int rc = annie_hip_create(l.data(), l.size(), a, device, min_ln, &h_);
    if (rc < 0) throw std::runtime_error(annie_hip_last_error());
    failed_ = (rc == ANNIE_INIT_FAILED);
    annie_hip_info(h_, &K_, &pck_, &LN_, &lbase_);
  }
  ~transform_hip() override { annie_hip_destroy(h_); }

  void set1() override { check(annie_hip_set1(h_)); }
  void get(mpz_t z, size_t src) override {
    std::vector<uint64_t> out(nlimbs_);
    size_t n = 0;
    check(annie_hip_get(h_, src, out.data(), out.size(), &n));
    from_limbs(z, out, n);
  }
  void copy(size_t dst, size_t src) override { check(annie_hip_copy(h_, dst, src)); }
  void square_mul(int32_t a) override { check(annie_hip_square_mul(h_, a)); }
  void square_mul_bits(const std::vector<uint8_t>& bits, int32_t a) override {
    check(annie_hip_square_mul_bits(h_, bits.data(), bits.size(), a));
  }
  void mul(size_t dst, size_t src) override { check(annie_hip_mul(h_, dst, src)); }
  bool init_failed() const override { return failed_; }
  void get_factor(mpz_t z) const override {
    std::vector<uint64_t> out(nlimbs_);
    size_t n = 0;
    check(annie_hip_get_factor(h_, out.data(), out.size(), &n));
    from_limbs(z, out, n);
  }
  bool has_checkpoint() const override { return true; }
  bool read_checkpoint(const std::vector<char>& payload) override {
    return annie_hip_load_checkpoint(h_, payload.data(), payload.size()) == ANNIE_OK;
  }
  std::vector<char> save_checkpoint() override {
    std::vector<char> buf(annie_hip_checkpoint_size(h_));
    check(annie_hip_save_checkpoint(h_, buf.data(), buf.size()));
    return buf;
  }
  std::string get_info() const override {
    std::ostringstream ss;
    ss << "Montgomery-McLaughlin-Gallot-Woltman reduction, " << K_ << "-prime NTT of length " << (size_t(1) << LN_)
       << ", packed-base = 2^{" << pck_ << "x" << lbase_ << "}";
    return ss.str();
  }
  void corrupt(size_t reg) override {
    // Replace one residue of the register by a different value in [0, p).
    std::vector<uint32_t> w(size_t(K_) << LN_);
    check(annie_hip_read_register(h_, reg, w.data(), w.size()));
    w[size_t(1) << (LN_ - 1)] ^= 1u;
    check(annie_hip_write_register(h_, reg, w.data(), w.size()));
  }
};

}  // namespace

transform* create_transform_gmp(const mpz_t N, int32_t a) { return new transform_gmp(N, a); }

transform* create_transform_hip(const mpz_t N, int32_t a, int device, int min_ln) {
  return new transform_hip(N, a, device, min_ln);
}

}  // namespace annie_host
