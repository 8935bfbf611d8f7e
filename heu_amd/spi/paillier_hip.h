// paillier_hip.h -- HEU vectorized-SPI classes over the engine's C ABI.
//
// The class set and method shapes follow HEU's backend contract:
//   heu/library/algorithms/clean_template/vector_encryptor.h:24-33
//   heu/library/algorithms/clean_template/vector_decryptor.h:25-31
//   heu/library/algorithms/clean_template/vector_evaluator.h:24-69
// and the key classes of heu/library/algorithms/paillier_zahlen
// (public_key.h:27-55, secret_key.h:22-55, key_generator.cc:26-60).
// ConstSpan<T> / Span<T> are spans of borrowed pointers, as in
// heu/library/algorithms/util/spi_traits.h:22-26.  Every batch call runs one
// engine entry point on the device; errors throw std::runtime_error (HEU
// throws yacl::Exception).
//
// Methods are const and re-entrant: HEU calls them from its worker threads,
// one chunk per thread (numpy/encryptor.cc:28, numpy/evaluator.cc:124), and
// scalar calls arrive as length-1 spans.  Elementwise calls coalesce: a call
// that finds the device busy queues, and the next thread to take the device
// runs every queued call of the same operation and key as one engine call.
#pragma once

#include <openssl/bn.h>

#include <cstdint>
#include <memory>
#include <span>
#include <string>
#include <utility>
#include <vector>

namespace heu_amd {

// Minimal arbitrary-precision signed integer (OpenSSL BIGNUM), enough for
// plaintexts and key material.
class BigInt {
 public:
  BigInt();
  BigInt(int64_t v);  // NOLINT: implicit like yacl::math::BigInt
  BigInt(const BigInt& o);
  BigInt(BigInt&& o) noexcept;
  BigInt& operator=(const BigInt& o);
  BigInt& operator=(BigInt&& o) noexcept;
  ~BigInt();

  static BigInt FromString(const std::string& s, int base = 10);
  static BigInt FromWords(const uint32_t* w, size_t words, bool twos_complement);
  // two's complement over `words` little-endian limbs; throws if it does not fit
  void ToWords(uint32_t* w, size_t words) const;

  std::string ToString() const;
  std::string ToHexString() const;
  size_t BitCount() const;
  bool IsNegative() const;
  bool IsZero() const;
  int Compare(const BigInt& o) const;
  int CompareAbs(const BigInt& o) const;
  bool FitsInt64() const;
  int64_t ToInt64() const;

  BigInt operator+(const BigInt& o) const;
  BigInt operator-(const BigInt& o) const;
  BigInt operator*(const BigInt& o) const;
  BigInt operator-() const;
  BigInt operator>>(int bits) const;
  bool operator==(const BigInt& o) const { return Compare(o) == 0; }
  bool operator!=(const BigInt& o) const { return Compare(o) != 0; }
  bool operator<(const BigInt& o) const { return Compare(o) < 0; }
  bool operator>(const BigInt& o) const { return Compare(o) > 0; }

  const BIGNUM* bn() const { return b_; }
  BIGNUM* bn() { return b_; }

 private:
  BIGNUM* b_;
};

template <class T>
using ConstSpan = std::span<const T* const>;
template <class T>
using Span = std::span<T* const>;

namespace paillier_hip {

using Plaintext = BigInt;
class PublicKey;

// Z-space value c = (1 + n m) h_s^r mod n^2 as little-endian uint32 limbs.
class Ciphertext {
 public:
  Ciphertext() = default;
  explicit Ciphertext(std::vector<uint32_t> limbs) : c_(std::move(limbs)) {}
  bool operator==(const Ciphertext& o) const { return c_ == o.c_; }
  bool operator!=(const Ciphertext& o) const { return c_ != o.c_; }
  std::string ToString() const;
  // paillier_z::Ciphertext's msgpack (MSGPACK_DEFINE(c_), paillier_zahlen/
  // ciphertext.h:37) with c_ in the Montgomery form the reference stores, c R
  // mod n^2 for R = 2^(64 ceil(bits(n^2) / 64)).  The key gives n^2, since c_
  // here is the Z-space value.  HEU's variant wrapper appends the schema index.
  std::vector<uint8_t> Serialize(const PublicKey& pk) const;
  static Ciphertext Deserialize(const PublicKey& pk, std::span<const uint8_t> in);
  std::vector<uint32_t> c_;
};

struct DeviceKey;  // engine handles (context + resident key)

class PublicKey {
 public:
  BigInt n_;         // n = p q
  BigInt n_square_;  // n^2
  BigInt n_half_;    // n / 2, the plaintext bound
  BigInt h_s_;       // h^n mod n^2
  size_t key_size_ = 0;
  std::shared_ptr<DeviceKey> dev_;  // public part resident on the device

  void Init();  // derives n^2, n/2 and admits the key to the device
  const BigInt& PlaintextBound() const { return n_half_; }
  size_t ct_words() const;
  size_t n_words() const;
  bool operator==(const PublicKey& o) const { return n_ == o.n_ && h_s_ == o.h_s_; }
  std::string ToString() const;
  // msgpack [n, h_s] (public_key.h:64-76); Deserialize calls Init(), as the
  // reference's msgpack convert does.
  std::vector<uint8_t> Serialize() const;
  void Deserialize(std::span<const uint8_t> in);
};

class SecretKey {
 public:
  BigInt lambda_, mu_, p_, q_;
  std::shared_ptr<DeviceKey> dev_;  // full key (with the CRT halves) on the device

  void Init(const PublicKey& pk);
  bool operator==(const SecretKey& o) const { return p_ == o.p_ && q_ == o.q_; }
  std::string ToString() const;
  // msgpack [lambda, mu, p, q] (secret_key.h:64-80); after Deserialize, call
  // Init(pk) to make the key resident (lambda and mu are checked against p, q).
  std::vector<uint8_t> Serialize() const;
  void Deserialize(std::span<const uint8_t> in);
};

class KeyGenerator {
 public:
  static void Generate(size_t key_size, SecretKey* sk, PublicKey* pk);
  static void Generate(SecretKey* sk, PublicKey* pk) { Generate(2048, sk, pk); }
  // reproducible keys (the engine's seeded counter stream)
  static void GenerateSeeded(size_t key_size, uint64_t seed, SecretKey* sk, PublicKey* pk);
};

class Encryptor {
 public:
  explicit Encryptor(const PublicKey& pk) : pk_(pk) {}
  std::vector<Ciphertext> EncryptZero(int64_t size) const;
  std::vector<Ciphertext> Encrypt(ConstSpan<Plaintext> pts) const;
  std::pair<std::vector<Ciphertext>, std::vector<std::string>> EncryptWithAudit(ConstSpan<Plaintext> pts) const;
  // deterministic variant: r given as r_words limbs per plaintext
  std::vector<Ciphertext> EncryptWithR(ConstSpan<Plaintext> pts, const uint32_t* r, size_t r_words) const;

  // single-element forms of the scalar SPI (paillier_zahlen/encryptor.h:25-37)
  Ciphertext EncryptZero() const;
  Ciphertext Encrypt(const Plaintext& m) const;
  std::pair<Ciphertext, std::string> EncryptWithAudit(const Plaintext& m) const;

 private:
  PublicKey pk_;
};

class Decryptor {
 public:
  Decryptor(const PublicKey& pk, const SecretKey& sk);
  std::vector<Plaintext> Decrypt(ConstSpan<Ciphertext> cts) const;
  void Decrypt(ConstSpan<Ciphertext> in_cts, Span<Plaintext> out_pts) const;
  // paillier_zahlen/decryptor.h:27-29
  void Decrypt(const Ciphertext& ct, Plaintext* out) const;
  Plaintext Decrypt(const Ciphertext& ct) const;

 private:
  PublicKey pk_;
  SecretKey sk_;
};

class Evaluator {
 public:
  explicit Evaluator(const PublicKey& pk) : pk_(pk) {}

  void Randomize(Span<Ciphertext> ct) const;

  // single-element forms (paillier_zahlen/evaluator.h:27-66)
  void Randomize(Ciphertext* ct) const;
  Ciphertext Add(const Ciphertext& a, const Ciphertext& b) const;
  Ciphertext Add(const Ciphertext& a, const Plaintext& b) const;
  void AddInplace(Ciphertext* a, const Ciphertext& b) const;
  void AddInplace(Ciphertext* a, const Plaintext& b) const;
  Ciphertext Sub(const Ciphertext& a, const Ciphertext& b) const;
  Ciphertext Sub(const Ciphertext& a, const Plaintext& b) const;
  Ciphertext Sub(const Plaintext& a, const Ciphertext& b) const;
  void SubInplace(Ciphertext* a, const Ciphertext& b) const;
  void SubInplace(Ciphertext* a, const Plaintext& b) const;
  Ciphertext Mul(const Ciphertext& a, const Plaintext& b) const;
  void MulInplace(Ciphertext* a, const Plaintext& b) const;
  Ciphertext Negate(const Ciphertext& a) const;
  void NegateInplace(Ciphertext* a) const;

  std::vector<Ciphertext> Add(ConstSpan<Ciphertext> a, ConstSpan<Ciphertext> b) const;
  std::vector<Ciphertext> Add(ConstSpan<Ciphertext> a, ConstSpan<Plaintext> b) const;
  std::vector<Ciphertext> Add(ConstSpan<Plaintext> a, ConstSpan<Ciphertext> b) const;
  std::vector<Plaintext> Add(ConstSpan<Plaintext> a, ConstSpan<Plaintext> b) const;
  void AddInplace(Span<Ciphertext> a, ConstSpan<Ciphertext> b) const;
  void AddInplace(Span<Ciphertext> a, ConstSpan<Plaintext> b) const;
  void AddInplace(Span<Plaintext> a, ConstSpan<Plaintext> b) const;

  std::vector<Ciphertext> Sub(ConstSpan<Ciphertext> a, ConstSpan<Ciphertext> b) const;
  std::vector<Ciphertext> Sub(ConstSpan<Ciphertext> a, ConstSpan<Plaintext> b) const;
  std::vector<Ciphertext> Sub(ConstSpan<Plaintext> a, ConstSpan<Ciphertext> b) const;
  std::vector<Plaintext> Sub(ConstSpan<Plaintext> a, ConstSpan<Plaintext> b) const;
  void SubInplace(Span<Ciphertext> a, ConstSpan<Ciphertext> b) const;
  void SubInplace(Span<Ciphertext> a, ConstSpan<Plaintext> p) const;
  void SubInplace(Span<Plaintext> a, ConstSpan<Plaintext> b) const;

  std::vector<Ciphertext> Mul(ConstSpan<Ciphertext> a, ConstSpan<Plaintext> b) const;
  std::vector<Ciphertext> Mul(ConstSpan<Plaintext> a, ConstSpan<Ciphertext> b) const;
  std::vector<Plaintext> Mul(ConstSpan<Plaintext> a, ConstSpan<Plaintext> b) const;
  void MulInplace(Span<Ciphertext> a, ConstSpan<Plaintext> b) const;
  void MulInplace(Span<Plaintext> a, ConstSpan<Plaintext> b) const;

  std::vector<Ciphertext> Negate(ConstSpan<Ciphertext> a) const;
  void NegateInplace(Span<Ciphertext> a) const;

  // sum of a span (the ReduceSum that numpy/evaluator.cc:102-104 looks for)
  Ciphertext ReduceSum(ConstSpan<Ciphertext> a) const;

 private:
  PublicKey pk_;
};

// Coalescing counters of this process: span calls submitted, engine calls that
// ran them, and the most calls one engine call served.
struct SpiStats {
  uint64_t calls = 0, device_calls = 0, max_calls_per_device_call = 0;
};
SpiStats GetSpiStats();
// Diagnostics: while held, queued calls wait (none takes the device), so a
// test can line up calls and release them into one engine call.
void HoldSpiQueue(bool hold);
// Calls waiting in the queue right now.
size_t SpiQueueLength();

}  // namespace paillier_hip
}  // namespace heu_amd
