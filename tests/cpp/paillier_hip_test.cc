// C++ tests of the vectorized-SPI classes (heu_amd/spi/paillier_hip.h) on a
// HIP device.  The first cases restate the reference's own unit tests,
// heu/library/algorithms/paillier_zahlen/paillier_test.cc:26-233
// (CiphertextEvaluate, MinMaxDecrypt, PlaintextEvaluate1/2, BigNumberTest
// EncDec/SubTest), through the single-element forms, and the reference's
// tests of its CUDA backend's vectorized SPI (paillier_gpu/paillier_gpu_test.cc)
// through every span overload, and the phe layer's evaluator and encryptor
// tests on this schema; the rest exercise the span (vectorized) forms
// against closed forms computed with OpenSSL on the host.  Prints one line per case and "ALL PASSED" at the end; exit status 0
// iff every check held.
#include <openssl/bn.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <limits>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "paillier_hip.h"

using heu_amd::BigInt;
using heu_amd::ConstSpan;
using heu_amd::Span;
using namespace heu_amd::paillier_hip;

namespace {

int g_failures = 0;

#define EXPECT_TRUE(c)                                                          \
  do {                                                                          \
    if (!(c)) {                                                                 \
      std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);              \
      ++g_failures;                                                             \
    }                                                                           \
  } while (0)
#define EXPECT_FALSE(c) EXPECT_TRUE(!(c))
#define EXPECT_EQ(a, b) EXPECT_TRUE((a) == (b))
#define EXPECT_THROW(stmt)                                                      \
  do {                                                                          \
    bool thrown_ = false;                                                       \
    try {                                                                       \
      stmt;                                                                     \
    } catch (const std::exception&) {                                           \
      thrown_ = true;                                                           \
    }                                                                           \
    if (!thrown_) {                                                             \
      std::printf("  FAILED %s:%d: no exception from %s\n", __FILE__, __LINE__, #stmt); \
      ++g_failures;                                                             \
    }                                                                           \
  } while (0)

struct Keys {
  SecretKey sk;
  PublicKey pk;
};

void run(const char* name, const std::function<void()>& body) {
  std::printf("[ RUN  ] %s\n", name);
  std::fflush(stdout);
  int before = g_failures;
  try {
    body();
  } catch (const std::exception& e) {
    std::printf("  FAILED: exception %s\n", e.what());
    ++g_failures;
  }
  std::printf("[ %s ] %s\n", g_failures == before ? " OK " : "FAIL", name);
  std::fflush(stdout);
}

template <class T>
std::vector<const T*> cptrs(const std::vector<T>& v) {
  std::vector<const T*> o;
  for (auto& x : v) o.push_back(&x);
  return o;
}

template <class T>
std::vector<T*> mptrs(std::vector<T>& v) {
  std::vector<T*> o;
  for (auto& x : v) o.push_back(&x);
  return o;
}

template <class T>
ConstSpan<T> cs(const std::vector<const T*>& v) {
  return ConstSpan<T>(v.data(), v.size());
}

template <class T>
Span<T> mspan(const std::vector<T*>& v) {
  return Span<T>(v.data(), v.size());
}

BigInt ct_value(const Ciphertext& c) { return BigInt::FromWords(c.c_.data(), c.c_.size(), false); }

BigInt mod_exp(const BigInt& b, const BigInt& e, const BigInt& m) {
  BigInt r;
  BN_CTX* ctx = BN_CTX_new();
  BN_mod_exp(r.bn(), b.bn(), e.bn(), m.bn(), ctx);
  BN_CTX_free(ctx);
  return r;
}

BigInt mod_mul(const BigInt& a, const BigInt& b, const BigInt& m) {
  BigInt r;
  BN_CTX* ctx = BN_CTX_new();
  BN_mod_mul(r.bn(), a.bn(), b.bn(), m.bn(), ctx);
  BN_CTX_free(ctx);
  return r;
}

BigInt nnmod(const BigInt& a, const BigInt& m) {
  BigInt r;
  BN_CTX* ctx = BN_CTX_new();
  BN_nnmod(r.bn(), a.bn(), m.bn(), ctx);
  BN_CTX_free(ctx);
  return r;
}

BigInt random_bits(std::mt19937_64& g, int bits, bool allow_negative) {
  std::vector<uint32_t> w((bits + 31) / 32);
  for (auto& x : w) x = (uint32_t)g();
  if (bits % 32) w.back() &= (1u << (bits % 32)) - 1u;
  BigInt v = BigInt::FromWords(w.data(), w.size(), false);
  return allow_negative && (g() & 1) ? -v : v;
}

// paillier_test.cc:26-96
void CiphertextEvaluate(const Keys& k) {
  Encryptor encryptor(k.pk);
  Evaluator evaluator(k.pk);
  Decryptor decryptor(k.pk, k.sk);
  EXPECT_TRUE(ct_value(encryptor.EncryptZero()) > BigInt(1));

  BigInt m0(-12345);
  Ciphertext ct0 = encryptor.Encrypt(m0);
  BigInt plain;
  Ciphertext res = evaluator.Add(ct0, ct0);
  decryptor.Decrypt(res, &plain);
  EXPECT_EQ(plain, BigInt(-12345 * 2));
  res = evaluator.Mul(ct0, BigInt(2));
  decryptor.Decrypt(res, &plain);
  EXPECT_EQ(plain, BigInt(-12345 * 2));
  evaluator.Randomize(&res);
  decryptor.Decrypt(res, &plain);
  EXPECT_EQ(plain, BigInt(-12345 * 2));

  BigInt m1(123);
  Ciphertext ct1 = encryptor.Encrypt(m1);
  res = evaluator.Add(ct1, ct1);
  decryptor.Decrypt(res, &plain);
  EXPECT_EQ(plain, BigInt(123 * 2));
  res = evaluator.Mul(ct1, BigInt(2));
  decryptor.Decrypt(res, &plain);
  EXPECT_EQ(plain, BigInt(123 * 2));
  res = evaluator.Add(ct0, ct1);
  decryptor.Decrypt(res, &plain);
  EXPECT_EQ(plain, BigInt(-12345 + 123));

  ct1 = encryptor.Encrypt(m1);
  res = evaluator.Mul(ct1, BigInt(1));
  decryptor.Decrypt(res, &plain);
  EXPECT_EQ(plain, BigInt(123));

  ct1 = encryptor.Encrypt(m1);
  res = evaluator.Mul(ct1, BigInt(0));
  decryptor.Decrypt(res, &plain);
  EXPECT_TRUE(plain.IsZero());
  evaluator.Randomize(&res);
  EXPECT_FALSE(ct_value(res).IsZero());
  decryptor.Decrypt(res, &plain);
  EXPECT_TRUE(plain.IsZero());

  ct1 = encryptor.Encrypt(m1);
  res = evaluator.Mul(ct1, BigInt(-1));
  decryptor.Decrypt(res, &plain);
  EXPECT_EQ(plain, BigInt(-123));

  ct1 = encryptor.Encrypt(m1);
  res = evaluator.Mul(ct1, BigInt(-2));
  decryptor.Decrypt(res, &plain);
  EXPECT_EQ(plain, BigInt(-123 * 2));
}

// paillier_test.cc:98-120
void MinMaxDecrypt(const Keys& k) {
  Encryptor encryptor(k.pk);
  Decryptor decryptor(k.pk, k.sk);
  BigInt plain = k.pk.n_;
  EXPECT_THROW(encryptor.Encrypt(plain));
  plain = k.pk.PlaintextBound() + BigInt(1);
  EXPECT_THROW(encryptor.Encrypt(plain));

  BigInt plain2;
  plain = plain - BigInt(1);  // max
  Ciphertext ct0 = encryptor.Encrypt(plain);
  decryptor.Decrypt(ct0, &plain2);
  EXPECT_EQ(plain, plain2);

  plain = -plain;  // -max
  ct0 = encryptor.Encrypt(plain);
  decryptor.Decrypt(ct0, &plain2);
  EXPECT_EQ(plain, plain2);

  plain = plain - BigInt(1);
  EXPECT_THROW(encryptor.Encrypt(plain));
}

// paillier_test.cc:122-186
void PlaintextEvaluate(const Keys& k) {
  Encryptor encryptor(k.pk);
  Evaluator evaluator(k.pk);
  Decryptor decryptor(k.pk, k.sk);
  const int64_t kMax = std::numeric_limits<int64_t>::max(), kMin = std::numeric_limits<int64_t>::lowest();

  Ciphertext ct0 = encryptor.Encrypt(BigInt(123));
  EXPECT_EQ(decryptor.Decrypt(evaluator.Add(ct0, BigInt(23))), BigInt(123 + 23));
  EXPECT_EQ(decryptor.Decrypt(evaluator.Add(ct0, BigInt(6543212))), BigInt(123 + 6543212));
  EXPECT_TRUE(decryptor.Decrypt(evaluator.Add(ct0, BigInt(-123))).IsZero());
  EXPECT_EQ(decryptor.Decrypt(evaluator.Add(ct0, BigInt(-456))), BigInt(123 - 456));

  ct0 = encryptor.Encrypt(BigInt(-123));
  EXPECT_EQ(decryptor.Decrypt(evaluator.Add(ct0, BigInt(23))), BigInt(-123 + 23));
  EXPECT_EQ(decryptor.Decrypt(evaluator.Add(ct0, BigInt(kMax))), BigInt(-123 + kMax));
  EXPECT_EQ(decryptor.Decrypt(evaluator.Add(ct0, BigInt(-123))), BigInt(-123 * 2));
  EXPECT_EQ(decryptor.Decrypt(evaluator.Add(ct0, BigInt(-456))), BigInt(-123 - 456));

  ct0 = encryptor.Encrypt(BigInt(kMin));
  EXPECT_EQ(decryptor.Decrypt(evaluator.Add(ct0, BigInt(kMax))), BigInt(-1));
  // kMin - 1 no longer fits int64 but is an ordinary plaintext
  BigInt below = decryptor.Decrypt(evaluator.Add(ct0, BigInt(-1)));
  EXPECT_EQ(below, BigInt(kMin) - BigInt(1));
  EXPECT_FALSE(below.FitsInt64());
}

// R mod n^2 of the Montgomery form the reference stores and prints, for
// yacl's 64-bit-word backends: R = 2^(64 ceil(bits(n^2) / 64)).
BigInt mont_r(const BigInt& n2) {
  BigInt R;
  BN_set_bit(R.bn(), (int)(64 * ((n2.BitCount() + 63) / 64)));
  return nnmod(R, n2);
}

// paillier_test.cc:188-233 (BigNumberTest EncDec / SubTest over the listed values)
void BigNumberEncDecSub(const Keys& k) {
  Encryptor encryptor(k.pk);
  Evaluator evaluator(k.pk);
  Decryptor decryptor(k.pk, k.sk);
  const int64_t vals[] = {std::numeric_limits<int64_t>::lowest(), -1234, -2, -1, 0, 1, 2, 1234,
                          std::numeric_limits<int64_t>::max()};
  for (int64_t x : vals) {
    BigInt pt1(x);
    EXPECT_EQ(decryptor.Decrypt(encryptor.Encrypt(pt1)), pt1);

    const int64_t share_a = std::numeric_limits<int64_t>::max();
    Ciphertext xe = encryptor.Encrypt(BigInt(x));
    auto xa = encryptor.EncryptWithAudit(BigInt(x));
    EXPECT_EQ(decryptor.Decrypt(xa.first), BigInt(x));
    const BigInt n2 = k.pk.n_ * k.pk.n_;
    const std::string c_hex = mod_mul(ct_value(xa.first), mont_r(n2), n2).ToHexString();
    EXPECT_TRUE(xa.second.rfind("p:", 0) == 0 && xa.second.find(",rn:") != std::string::npos &&
                xa.second.find(",c:" + c_hex) != std::string::npos);
    evaluator.SubInplace(&xe, BigInt(share_a));
    BigInt raw = decryptor.Decrypt(xe);
    // share_a + share_b == x in int64 arithmetic (wrapping, as the reference's Get<int64_t>)
    uint64_t share_b = raw.FitsInt64() ? (uint64_t)raw.ToInt64() : 0;
    if (!raw.FitsInt64()) {
      uint32_t w[4];
      raw.ToWords(w, 4);
      share_b = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
    }
    EXPECT_EQ((int64_t)((uint64_t)share_a + share_b), x);
  }
}

// Span forms: EncryptWithR against (1 + n m) h_s^r mod n^2 computed on the
// host, then every evaluator op against its closed form / decryption.
void VectorizedClosedForms(const Keys& k, size_t count) {
  Encryptor encryptor(k.pk);
  Evaluator evaluator(k.pk);
  Decryptor decryptor(k.pk, k.sk);
  const BigInt& n = k.pk.n_;
  const BigInt& n2 = k.pk.n_square_;
  std::mt19937_64 g(42 + count);
  const int nbits = (int)n.BitCount();

  std::vector<BigInt> ms, ms2, ks;
  for (size_t i = 0; i < count; ++i) {
    int bits = (int)(g() % (uint64_t)(nbits - 2)) + 1;
    ms.push_back(i % 3 == 0 ? BigInt((int64_t)g()) : random_bits(g, bits, true));
    ms2.push_back(BigInt((int64_t)(g() >> 1)) - BigInt((int64_t)(g() >> 1)));
    ks.push_back(i % 2 ? BigInt((int64_t)(int32_t)g()) : random_bits(g, (int)(g() % 300) + 1, true));
  }
  ms[0] = k.pk.n_half_;
  ms[1] = -k.pk.n_half_;
  ms[2] = BigInt(0);
  ks[0] = BigInt(0);
  ks[1] = BigInt(1);
  ks[2] = BigInt(-1);

  const size_t rw = k.pk.key_size_ / 64;
  std::vector<uint32_t> r(count * rw);
  for (auto& x : r) x = (uint32_t)g();
  auto pm = cptrs(ms), pm2 = cptrs(ms2), pk = cptrs(ks);
  auto ct = encryptor.EncryptWithR(cs(pm), r.data(), rw);
  bool enc_ok = true;
  for (size_t i = 0; i < count && enc_ok; ++i) {
    BigInt ri = BigInt::FromWords(r.data() + i * rw, rw, false);
    BigInt gm = nnmod(n * ms[i] + BigInt(1), n2);
    enc_ok = ct_value(ct[i]) == mod_mul(gm, mod_exp(k.pk.h_s_, ri, n2), n2);
  }
  EXPECT_TRUE(enc_ok);

  auto pc = cptrs(ct);
  auto dec = decryptor.Decrypt(cs(pc));
  bool dec_ok = true;
  for (size_t i = 0; i < count; ++i) dec_ok = dec_ok && dec[i] == ms[i];
  EXPECT_TRUE(dec_ok);

  // ct x big scalar: c^k mod n^2 (inverse power for k < 0)
  auto prod = evaluator.Mul(cs(pc), cs(pk));
  bool mul_ok = true;
  for (size_t i = 0; i < count && mul_ok; ++i) {
    BigInt e = ks[i].IsNegative() ? -ks[i] : ks[i];
    BigInt want = mod_exp(ct_value(ct[i]), e, n2);
    if (ks[i].IsNegative()) BN_mod_inverse(want.bn(), want.bn(), n2.bn(), nullptr);
    mul_ok = ct_value(prod[i]) == want;
  }
  EXPECT_TRUE(mul_ok);

  // ct + int64 pt, ct - pt, pt - ct, ct + ct, negate: decrypt to the plaintext arithmetic
  auto ct2 = encryptor.Encrypt(cs(pm2));
  auto pc2 = cptrs(ct2);
  auto dsum = decryptor.Decrypt(cs(cptrs(evaluator.Add(cs(pc), cs(pc2)))));
  auto dadd = decryptor.Decrypt(cs(cptrs(evaluator.Add(cs(pc), cs(pm2)))));
  auto dsub = decryptor.Decrypt(cs(cptrs(evaluator.Sub(cs(pc), cs(pc2)))));
  auto dsubp = decryptor.Decrypt(cs(cptrs(evaluator.Sub(cs(pm2), cs(pc)))));
  auto dneg = decryptor.Decrypt(cs(cptrs(evaluator.Negate(cs(pc)))));
  auto wrap = [&](const BigInt& v) {  // centered residue mod n
    BigInt m = nnmod(v, n);
    return m > k.pk.n_half_ ? m - n : m;
  };
  bool arith_ok = true;
  for (size_t i = 0; i < count; ++i) {
    arith_ok = arith_ok && dsum[i] == wrap(ms[i] + ms2[i]) && dadd[i] == wrap(ms[i] + ms2[i]) &&
               dsub[i] == wrap(ms[i] - ms2[i]) && dsubp[i] == wrap(ms2[i] - ms[i]) && dneg[i] == wrap(-ms[i]);
  }
  EXPECT_TRUE(arith_ok);

  // ct + big pt goes through Enc(p, r = 0)
  auto dbig = decryptor.Decrypt(cs(cptrs(evaluator.Add(cs(pc), cs(pm)))));
  bool big_ok = true;
  for (size_t i = 0; i < count; ++i) big_ok = big_ok && dbig[i] == wrap(ms[i] + ms[i]);
  EXPECT_TRUE(big_ok);

  // in-place forms and ReduceSum
  auto acc = ct2;
  auto macc = mptrs(acc);
  evaluator.AddInplace(mspan(macc), cs(pc));
  evaluator.MulInplace(mspan(macc), cs(pk));
  evaluator.Randomize(mspan(macc));
  auto dacc = decryptor.Decrypt(cs(cptrs(acc)));
  bool inplace_ok = true;
  for (size_t i = 0; i < count; ++i) inplace_ok = inplace_ok && dacc[i] == wrap((ms[i] + ms2[i]) * ks[i]);
  EXPECT_TRUE(inplace_ok);
  BigInt total(0);
  for (auto& v : ms2) total = total + v;
  EXPECT_EQ(decryptor.Decrypt(evaluator.ReduceSum(cs(pc2))), wrap(total));

  // EncryptZero batch decrypts to zeros
  auto z = encryptor.EncryptZero((int64_t)count);
  bool zero_ok = true;
  for (auto& p : decryptor.Decrypt(cs(cptrs(z)))) zero_ok = zero_ok && p.IsZero();
  EXPECT_TRUE(zero_ok);

  // an invalid ciphertext (>= n^2) is rejected like evaluator.cc:20-22
  Ciphertext bad = ct[0];
  std::vector<uint32_t> n2w(bad.c_.size() + 1);
  n2.ToWords(n2w.data(), n2w.size());
  bad.c_.assign(n2w.begin(), n2w.end() - 1);
  EXPECT_THROW(evaluator.Negate(bad));
  // span length mismatch
  std::vector<const Ciphertext*> shorter(pc.begin(), pc.end() - 1);
  EXPECT_THROW(evaluator.Add(cs(pc), cs(shorter)));
}

// The reference's tests of its CUDA backend's vectorized SPI,
// heu/library/algorithms/paillier_gpu/paillier_gpu_test.cc:117-763 and
// paillier_gpu_test_con.cc:53-105, restated
// with their values against the span forms: every Add / Sub / Mul overload
// (ct, ct / ct, pt / pt, ct / pt, pt), the in-place forms, Negate and
// Randomize.  Each test's value rows run as one span call.
void GpuBackendCases(const Keys& k) {
  Encryptor enc(k.pk);
  Evaluator ev(k.pk);
  Decryptor dec(k.pk, k.sk);
  const BigInt lo(std::numeric_limits<int64_t>::lowest()), hi(std::numeric_limits<int64_t>::max());
  using V = std::vector<BigInt>;
  auto E = [&](const V& v) { return enc.Encrypt(cs(cptrs(v))); };
  auto D = [&](const std::vector<Ciphertext>& c) { return dec.Decrypt(cs(cptrs(c))); };
  auto same = [](const V& a, const V& b) { return a == b; };

  // EncDecTest (:117-125)
  V pairs{1, 2, -1, 0, -3, hi, -65530, -123, lo, 0};
  EXPECT_TRUE(same(D(E(pairs)), pairs));
  // paillier_gpu_test_con.cc: EncDecBigintTest (:53-73) up to the plaintext
  // bound, and EncDecLongTest (:75-105), 100,000 elements of 33
  const BigInt top = k.pk.n_half_ - BigInt(1);
  V big{0, 1, 100, lo, top, -top};
  EXPECT_TRUE(same(D(E(big)), big));
  V many(100000, BigInt(33));
  EXPECT_TRUE(same(D(E(many)), many));
  // CiphertextsAddTest (:157-165)
  V a{-2, 123, -123, 234, lo, lo, hi}, b{-4, 234, 234, -123, hi, lo, hi};
  V c{-6, 357, 111, 111, -1, lo + lo, hi + hi};
  EXPECT_TRUE(same(D(ev.Add(cs(cptrs(E(a))), cs(cptrs(E(b))))), c));
  // CiphertextsAddInplaceTest (:192-197)
  {
    V x{5, 5, -5, -5}, y{2, -7, 7, -7}, z{7, -2, 2, -12};
    auto cx = E(x), cy = E(y);
    auto mx = mptrs(cx);
    ev.AddInplace(mspan(mx), cs(cptrs(cy)));
    EXPECT_TRUE(same(D(cx), z));
  }
  // CiphertextsRandomize (:217-243): same plaintext, another ciphertext
  {
    V x{5, 0, -123, lo, hi};
    auto cx = E(x);
    const auto before = cx;
    auto mx = mptrs(cx);
    ev.Randomize(mspan(mx));
    EXPECT_TRUE(same(D(cx), x));
    bool all_changed = true;
    for (size_t i = 0; i < cx.size(); ++i) all_changed = all_changed && cx[i] != before[i];
    EXPECT_TRUE(all_changed);
  }
  // PlaintextsAddTest / PlaintextsAddInplace (:245-302)
  {
    V x{-5, 7, 3, 7, -3, -7}, y{8, -9, 8, 9, -8, -9}, z{3, -2, 11, 16, -11, -16};
    EXPECT_TRUE(same(ev.Add(cs(cptrs(x)), cs(cptrs(y))), z));
    auto mx = mptrs(x);
    ev.AddInplace(mspan(mx), cs(cptrs(y)));
    EXPECT_TRUE(same(x, z));
  }
  // CiphertextsAddPlaintexts, PlaintextsAddCiphertexts(Inplace) (:304-401)
  {
    V x{123, 123, -123, -123, lo, hi}, y{2, -2, 2, -2, hi, -2}, z{125, 121, -121, -125, -1, hi - BigInt(2)};
    auto cx = E(x);
    EXPECT_TRUE(same(D(ev.Add(cs(cptrs(cx)), cs(cptrs(y)))), z));
    EXPECT_TRUE(same(D(ev.Add(cs(cptrs(y)), cs(cptrs(cx)))), z));
    auto mx = mptrs(cx);
    ev.AddInplace(mspan(mx), cs(cptrs(y)));
    EXPECT_TRUE(same(D(cx), z));
  }
  // CtsMulPtsTest, PtsMulCtsTest, CtsMulPtsInplaceTest (:403-508)
  {
    V x{2, -2, 2, -2}, y{4, 4, -4, -4}, z{8, -8, -8, 8};
    auto cx = E(x);
    EXPECT_TRUE(same(D(ev.Mul(cs(cptrs(cx)), cs(cptrs(y)))), z));
    EXPECT_TRUE(same(D(ev.Mul(cs(cptrs(y)), cs(cptrs(cx)))), z));
    auto mx = mptrs(cx);
    ev.MulInplace(mspan(mx), cs(cptrs(y)));
    EXPECT_TRUE(same(D(cx), z));
  }
  // PtsMulPtsTest, PtsMulPtsInplaceTest (:458-533)
  {
    V x{5, 7, -5, 7, -5, -7}, y{8, 9, 8, -9, -8, -9}, z{40, 63, -40, -63, 40, 63};
    EXPECT_TRUE(same(ev.Mul(cs(cptrs(x)), cs(cptrs(y))), z));
    auto mx = mptrs(x);
    ev.MulInplace(mspan(mx), cs(cptrs(y)));
    EXPECT_TRUE(same(x, z));
  }
  // CiphertextsNegateTest, CiphertextsNegateInplaceTest (:535-593)
  {
    V x{5, -123, 0, hi}, z{-5, 123, 0, -hi};
    auto cx = E(x);
    EXPECT_TRUE(same(D(ev.Negate(cs(cptrs(cx)))), z));
    auto mx = mptrs(cx);
    ev.NegateInplace(mspan(mx));
    EXPECT_TRUE(same(D(cx), z));
  }
  // CiphertextsSub, CtsSubPts, PtsSubCts, PtsSubPts and their in-place forms
  // (:595-763): a - b with a, b in {(123, -7), (-123, -7), (123, 7)}, plus
  // (iLow, -1) for ct - ct
  {
    V x{123, -123, 123, lo}, y{-7, -7, 7, -1}, z{130, -116, 116, lo + BigInt(1)};
    auto cx = E(x), cy = E(y);
    EXPECT_TRUE(same(D(ev.Sub(cs(cptrs(cx)), cs(cptrs(cy)))), z));
    EXPECT_TRUE(same(D(ev.Sub(cs(cptrs(cx)), cs(cptrs(y)))), z));
    EXPECT_TRUE(same(D(ev.Sub(cs(cptrs(x)), cs(cptrs(cy)))), z));
    EXPECT_TRUE(same(ev.Sub(cs(cptrs(x)), cs(cptrs(y))), z));
    auto cx2 = cx;
    auto m1 = mptrs(cx), m2 = mptrs(cx2);
    ev.SubInplace(mspan(m1), cs(cptrs(cy)));
    ev.SubInplace(mspan(m2), cs(cptrs(y)));
    EXPECT_TRUE(same(D(cx), z));
    EXPECT_TRUE(same(D(cx2), z));
    auto mx = mptrs(x);
    ev.SubInplace(mspan(mx), cs(cptrs(y)));
    EXPECT_TRUE(same(x, z));
  }
}

// The phe layer's tests on the ZPaillier schema, through the single-element
// forms: heu/library/phe/evaluator_test.cc:31-140 (Evaluate, EvaluateInplace)
// and encryptor_test.cc:32-90 (EncryptZero, MinMaxEnc).  Plaintexts are what
// PlainEncoder makes of them: Encode(x) = x * scale, with scale 1 (`edr`) or
// INT64_MAX (`encoder`, beyond int64).
void PheCases(const Keys& k) {
  Encryptor enc(k.pk);
  Evaluator ev(k.pk);
  Decryptor dec(k.pk, k.sk);
  // Evaluate
  Ciphertext ct0 = enc.Encrypt(BigInt(-12345));
  ev.Randomize(&ct0);
  EXPECT_EQ(dec.Decrypt(ct0), BigInt(-12345));
  EXPECT_EQ(dec.Decrypt(ev.Negate(ct0)), BigInt(12345));
  EXPECT_EQ(dec.Decrypt(ev.Add(ct0, ct0)), BigInt(-12345 * 2));
  EXPECT_EQ(dec.Decrypt(ev.Add(ct0, BigInt(345))), BigInt(-12345 + 345));
  const BigInt scale(std::numeric_limits<int64_t>::max());
  const BigInt imax = scale;
  const Ciphertext ct_max = enc.Encrypt((imax - BigInt(1)) * scale);
  EXPECT_EQ(dec.Decrypt(ev.Add(ct_max, BigInt(1) * scale)), imax * scale);
  EXPECT_EQ(dec.Decrypt(ev.Sub(ct0, ct0)), BigInt(0));
  EXPECT_EQ(dec.Decrypt(ev.Sub(ct0, BigInt(345))), BigInt(-12345 - 345));
  EXPECT_EQ(dec.Decrypt(ev.Sub(BigInt(789), ct0)), BigInt(12345 + 789));
  const BigInt umax = imax * BigInt(2) + BigInt(1);  // UINT64_MAX
  const Ciphertext ct_umax = enc.Encrypt(umax * scale);
  EXPECT_EQ(dec.Decrypt(ev.Sub(ct_umax, BigInt(1) * scale)), (umax - BigInt(1)) * scale);
  for (int64_t m : {0, 1, 2, 10}) EXPECT_EQ(dec.Decrypt(ev.Mul(ct0, BigInt(m))), BigInt(-12345 * m));
  // EvaluateInplace
  BigInt pt0(9876), pt1(1234);
  Ciphertext c0 = enc.Encrypt(pt0), c1 = enc.Encrypt(pt1);
  ev.Randomize(&c0);
  EXPECT_EQ(dec.Decrypt(c0), BigInt(9876));
  ev.NegateInplace(&c1);
  pt1 = -pt1;
  EXPECT_EQ(dec.Decrypt(c1), pt1);
  ev.AddInplace(&c0, c1);
  pt0 = pt0 + pt1;
  EXPECT_EQ(dec.Decrypt(c0), pt0);
  ev.AddInplace(&c0, BigInt(963));
  pt0 = pt0 + BigInt(963);
  EXPECT_EQ(dec.Decrypt(c0), pt0);
  ev.AddInplace(&c0, BigInt(-741));
  pt0 = pt0 + BigInt(-741);
  EXPECT_EQ(dec.Decrypt(c0), pt0);
  ev.SubInplace(&c0, c1);
  pt0 = pt0 - pt1;
  EXPECT_EQ(dec.Decrypt(c0), pt0);
  ev.SubInplace(&c0, BigInt(852));
  pt0 = pt0 - BigInt(852);
  EXPECT_EQ(dec.Decrypt(c0), pt0);
  ev.MulInplace(&c0, BigInt(10));
  pt0 = pt0 * BigInt(10);
  EXPECT_EQ(dec.Decrypt(c0), pt0);
  // EncryptZero
  Ciphertext z = enc.EncryptZero();
  EXPECT_EQ(dec.Decrypt(z), BigInt(0));
  ev.AddInplace(&z, enc.EncryptZero());
  EXPECT_EQ(dec.Decrypt(z), BigInt(0));
  ev.SubInplace(&z, enc.EncryptZero());
  EXPECT_EQ(dec.Decrypt(z), BigInt(0));
  EXPECT_EQ(dec.Decrypt(ev.Sub(BigInt(123), z)), BigInt(123));
  for (int64_t m : {0, 123456, -123456}) {
    ev.MulInplace(&z, BigInt(m));
    EXPECT_EQ(dec.Decrypt(z), BigInt(0));
  }
  ev.NegateInplace(&z);
  EXPECT_EQ(dec.Decrypt(z), BigInt(0));
  // MinMaxEnc: |m| <= PlaintextBound (n / 2) encrypts, one past it throws
  const BigInt bound = k.pk.n_half_;
  EXPECT_THROW(enc.Encrypt(bound + BigInt(1)));
  EXPECT_THROW(enc.Encrypt(-(bound + BigInt(1))));
  EXPECT_EQ(dec.Decrypt(enc.Encrypt(bound)), bound);
  EXPECT_EQ(dec.Decrypt(enc.Encrypt(-bound)), -bound);
}

// Spans long enough that the SPI splits its host work (packing, checks,
// conversions, result objects) over threads: every path still gives the
// closed forms, checked through decryption, and an error names the first bad
// element in span order, as a serial loop would.
void SpansSplitOverThreads(const Keys& k) {
  // 50,000 elements: split over threads on the host and into three pipelined
  // pieces of about 16,667 on the device
  const size_t count = 50000;
  Encryptor encryptor(k.pk);
  Evaluator evaluator(k.pk);
  Decryptor decryptor(k.pk, k.sk);
  const BigInt& n = k.pk.n_;
  const int nbits = (int)n.BitCount();
  std::mt19937_64 g(7);
  std::vector<BigInt> big, small, k64, kbig;
  for (size_t i = 0; i < count; ++i) {
    big.push_back(random_bits(g, (int)(g() % (uint64_t)(nbits - 2)) + 1, true));
    small.push_back(BigInt((int64_t)g()));
    k64.push_back(BigInt((int64_t)(int32_t)g()));
    kbig.push_back(random_bits(g, 100, true));
  }
  auto wrap = [&](const BigInt& v) {
    BigInt m = nnmod(v, n);
    return m > k.pk.n_half_ ? m - n : m;
  };
  auto pb = cptrs(big), ps = cptrs(small), pk64 = cptrs(k64), pkb = cptrs(kbig);
  auto cb = encryptor.Encrypt(cs(pb));  // plaintexts beyond int64
  auto csm = encryptor.Encrypt(cs(ps));
  auto pcb = cptrs(cb), pcs = cptrs(csm);
  auto db = decryptor.Decrypt(cs(pcb)), dsm = decryptor.Decrypt(cs(pcs));
  auto dm64 = decryptor.Decrypt(cs(cptrs(evaluator.Mul(cs(pcs), cs(pk64)))));
  auto dmbig = decryptor.Decrypt(cs(cptrs(evaluator.Mul(cs(pcs), cs(pkb)))));
  auto dadd = decryptor.Decrypt(cs(cptrs(evaluator.Add(cs(pcb), cs(ps)))));
  auto dsub = decryptor.Decrypt(cs(cptrs(evaluator.Sub(cs(pcb), cs(pcs)))));
  bool ok = db.size() == count && dsm.size() == count;
  for (size_t i = 0; ok && i < count; ++i)
    ok = db[i] == big[i] && dsm[i] == small[i] && dm64[i] == wrap(small[i] * k64[i]) &&
         dmbig[i] == wrap(small[i] * kbig[i]) && dadd[i] == wrap(big[i] + small[i]) &&
         dsub[i] == wrap(big[i] - small[i]);
  EXPECT_TRUE(ok);

  // out-of-range plaintexts at 4100 (past the first host chunk) and 45000 (in
  // the last piece): the message names the one at 4100; then at 20000 (the
  // second piece, after the first ran) and 45000: it names the one at 20000
  auto names_first = [&](const std::function<void()>& call, const BigInt& first) {
    try {
      call();
    } catch (const std::exception& e) {
      return std::string(e.what()).find("message=" + first.ToHexString() + ",") != std::string::npos;
    }
    return false;
  };
  std::vector<BigInt> bad = small;
  bad[4100] = k.pk.n_half_ + BigInt(7);
  bad[45000] = k.pk.n_half_ + BigInt(9);
  auto pbad = cptrs(bad);
  EXPECT_TRUE(names_first([&] { encryptor.Encrypt(cs(pbad)); }, bad[4100]));
  EXPECT_TRUE(names_first([&] { evaluator.Add(cs(pcs), cs(pbad)); }, bad[4100]));
  bad[4100] = small[4100];
  bad[20000] = k.pk.n_half_ + BigInt(11);
  EXPECT_TRUE(names_first([&] { encryptor.Encrypt(cs(pbad)); }, bad[20000]));
  EXPECT_TRUE(names_first([&] { evaluator.Add(cs(pcs), cs(pbad)); }, bad[20000]));
  EXPECT_TRUE(names_first([&] { evaluator.Sub(cs(pcs), cs(pbad)); }, bad[20000]));
  // a ciphertext >= n^2 deep in the span (the third piece) is rejected, and the
  // calls after it still give every element right
  auto cbad = csm;
  std::vector<uint32_t> n2w(cbad[0].c_.size() + 1);
  k.pk.n_square_.ToWords(n2w.data(), n2w.size());
  cbad[40000].c_.assign(n2w.begin(), n2w.end() - 1);
  EXPECT_THROW(evaluator.Add(cs(cptrs(cbad)), cs(pcs)));
  EXPECT_THROW(evaluator.Negate(cs(cptrs(cbad))));
  EXPECT_THROW(decryptor.Decrypt(cs(cptrs(cbad))));
  auto dagain = decryptor.Decrypt(cs(pcs));
  bool again = dagain.size() == count;
  for (size_t i = 0; again && i < count; ++i) again = dagain[i] == small[i];
  EXPECT_TRUE(again);
}

// HEU's numpy layer calls the span methods from its worker threads, one chunk
// per thread (numpy/encryptor.cc:28, numpy/evaluator.cc:124).  Calls lined up
// while the queue is held run as one engine call per (operation, key), and
// every thread gets what it gets alone.  One Mul chunk holds a ciphertext 0
// times -1 (no inverse): only that call throws.
void ConcurrentChunks(const Keys& k, const Keys& k2) {
  Encryptor enc(k.pk), enc2(k2.pk);
  Decryptor dec(k.pk, k.sk);
  Evaluator ev(k.pk), ev2(k2.pk);
  constexpr int T = 8;
  constexpr size_t kPer = 33;
  std::mt19937_64 g(7);
  std::vector<std::vector<BigInt>> m(T), w(T);
  std::vector<std::vector<Ciphertext>> c(T), c2(T);
  for (int t = 0; t < T; ++t) {
    for (size_t i = 0; i < kPer; ++i) {
      m[t].push_back(BigInt((int64_t)g()));
      w[t].push_back(BigInt((int64_t)(int32_t)g()));
    }
    c[t] = enc.Encrypt(cs(cptrs(m[t])));
    c2[t] = enc2.Encrypt(cs(cptrs(m[t])));
  }
  // the poisoned Mul chunk: element 5 is the ciphertext 0, times -1
  std::vector<Ciphertext> bad = c[3];
  bad[5].c_.assign(bad[5].c_.size(), 0u);
  std::vector<BigInt> wbad = w[3];
  wbad[5] = BigInt(-1);
  // one thread, one call at a time
  std::vector<std::vector<Ciphertext>> add_want(T), add2_want(T), mul_want(T);
  std::vector<std::vector<BigInt>> dec_want(T);
  for (int t = 0; t < T; ++t) {
    add_want[t] = ev.Add(cs(cptrs(c[t])), cs(cptrs(c[(t + 1) % T])));
    add2_want[t] = ev2.Add(cs(cptrs(c2[t])), cs(cptrs(c2[(t + 2) % T])));
    if (t != 3) mul_want[t] = ev.Mul(cs(cptrs(c[t])), cs(cptrs(w[t])));
    dec_want[t] = dec.Decrypt(cs(cptrs(c[t])));
  }
  std::vector<std::vector<Ciphertext>> add_got(T), add2_got(T), mul_got(T);
  std::vector<std::vector<BigInt>> dec_got(T);
  std::vector<int> threw(T, 0);
  const SpiStats s0 = GetSpiStats();
  HoldSpiQueue(true);
  std::vector<std::thread> th;
  for (int t = 0; t < T; ++t) {
    th.emplace_back([&, t] { add_got[t] = ev.Add(cs(cptrs(c[t])), cs(cptrs(c[(t + 1) % T]))); });
    th.emplace_back([&, t] { add2_got[t] = ev2.Add(cs(cptrs(c2[t])), cs(cptrs(c2[(t + 2) % T]))); });
    th.emplace_back([&, t] {
      try {
        mul_got[t] = t == 3 ? ev.Mul(cs(cptrs(bad)), cs(cptrs(wbad))) : ev.Mul(cs(cptrs(c[t])), cs(cptrs(w[t])));
      } catch (const std::exception&) {
        threw[t] = 1;
      }
    });
    th.emplace_back([&, t] { dec_got[t] = dec.Decrypt(cs(cptrs(c[t]))); });
  }
  const auto deadline = std::chrono::steady_clock::now() + std::chrono::seconds(30);
  while (SpiQueueLength() < (size_t)(4 * T) && std::chrono::steady_clock::now() < deadline)
    std::this_thread::sleep_for(std::chrono::milliseconds(1));
  EXPECT_EQ(SpiQueueLength(), (size_t)(4 * T));
  HoldSpiQueue(false);
  for (auto& x : th) x.join();
  const SpiStats s1 = GetSpiStats();
  std::printf("  %llu calls -> %llu engine calls (largest group %llu calls)\n",
              (unsigned long long)(s1.calls - s0.calls), (unsigned long long)(s1.device_calls - s0.device_calls),
              (unsigned long long)s1.max_calls_per_device_call);
  EXPECT_EQ(s1.calls - s0.calls, (uint64_t)(4 * T));
  // Add, Add on the second key, Decrypt: one call each; Mul: the failed group
  // call, then its 8 calls one by one (7 succeed)
  EXPECT_EQ(s1.device_calls - s0.device_calls, (uint64_t)(3 + T - 1));
  EXPECT_TRUE(s1.max_calls_per_device_call >= (uint64_t)T);
  for (int t = 0; t < T; ++t) {
    EXPECT_EQ(add_got[t].size(), kPer);
    for (size_t i = 0; i < add_got[t].size(); ++i) EXPECT_TRUE(add_got[t][i].c_ == add_want[t][i].c_);
    for (size_t i = 0; i < add2_got[t].size(); ++i) EXPECT_TRUE(add2_got[t][i].c_ == add2_want[t][i].c_);
    EXPECT_EQ(add2_got[t].size(), kPer);
    EXPECT_TRUE(dec_got[t] == dec_want[t]);
    EXPECT_TRUE(dec_got[t] == m[t]);
    EXPECT_EQ(threw[t], t == 3 ? 1 : 0);
    if (t != 3) {
      EXPECT_EQ(mul_got[t].size(), kPer);
      for (size_t i = 0; i < mul_got[t].size(); ++i) EXPECT_TRUE(mul_got[t][i].c_ == mul_want[t][i].c_);
    }
  }
}

}  // namespace

// Throughput through the span forms with host spans: what a HEU caller sees,
// including the host conversions, one H2D and one D2H copy and a sync per
// engine call.  Printed (the engine's own rates are measured on device-resident
// data elsewhere); the decryptions are checked against the plaintexts.
void SpanRates(const Keys& k, size_t count) {
  Encryptor enc(k.pk);
  Decryptor dec(k.pk, k.sk);
  Evaluator ev(k.pk);
  std::mt19937_64 g(11);
  std::vector<BigInt> m, w;
  for (size_t i = 0; i < count; ++i) {
    m.push_back(BigInt((int64_t)g()));
    w.push_back(BigInt((int64_t)(int32_t)g()));
  }
  auto secs = [](const std::function<void()>& f) {
    const auto t0 = std::chrono::steady_clock::now();
    f();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  std::vector<Ciphertext> c, s, p;
  std::vector<BigInt> d;
  c = enc.Encrypt(cs(cptrs(m)));  // warm-up: staging arena, scratch
  d = dec.Decrypt(cs(cptrs(c)));
  const double te = secs([&] { c = enc.Encrypt(cs(cptrs(m))); });
  const double ta = secs([&] { s = ev.Add(cs(cptrs(c)), cs(cptrs(c))); });
  const double tm = secs([&] { p = ev.Mul(cs(cptrs(c)), cs(cptrs(w))); });
  const double td = secs([&] { d = dec.Decrypt(cs(cptrs(c))); });
  bool same = d.size() == count;
  for (size_t i = 0; same && i < count; ++i) same = d[i] == m[i];
  EXPECT_TRUE(same);
  // every element of the sums and products as well
  const std::vector<BigInt> ds = dec.Decrypt(cs(cptrs(s))), dp = dec.Decrypt(cs(cptrs(p)));
  bool sums = ds.size() == count && dp.size() == count;
  for (size_t i = 0; sums && i < count; ++i) sums = ds[i] == m[i] + m[i] && dp[i] == m[i] * w[i];
  EXPECT_TRUE(sums);
  std::printf("  span rates, %zu elements (elements/s, host spans, PCIe included): encrypt %.0f, add %.0f, "
              "mul %.0f, decrypt %.0f\n",
              count, count / te, count / ta, count / tm, count / td);
}

// The backend objects' serialized form: msgpack of paillier_z::Ciphertext
// (ciphertext.h:37, c_ in Montgomery form), PublicKey [n, h_s] and SecretKey
// [lambda, mu, p, q] (public_key.h:64-76, secret_key.h:64-80); keys and
// ciphertexts re-created from the bytes decrypt as before, and the audit
// strings carry the Montgomery forms (encryptor.cc:50-54).
BigInt str_value(const std::vector<uint8_t>& b, size_t at) {
  size_t len = 0, off = 0;
  if (b[at] >= 0xa0 && b[at] <= 0xbf) len = b[at] & 0x1f, off = at + 1;
  else if (b[at] == 0xd9) len = b[at + 1], off = at + 2;
  else if (b[at] == 0xda) len = ((size_t)b[at + 1] << 8) | b[at + 2], off = at + 3;
  EXPECT_TRUE(len >= 1 && b[off] == 0);  // sign byte: non-negative
  std::vector<uint8_t> mag(b.begin() + off + 1, b.begin() + off + len);
  mag.resize((mag.size() + 3) / 4 * 4 + 4, 0);
  return BigInt::FromWords(reinterpret_cast<const uint32_t*>(mag.data()), mag.size() / 4, false);
}

void Serialization(const Keys& k) {
  const BigInt n2 = k.pk.n_ * k.pk.n_;
  const BigInt R = mont_r(n2);
  PublicKey pk2;
  pk2.Deserialize(k.pk.Serialize());
  EXPECT_TRUE(pk2 == k.pk);
  const auto skb = k.sk.Serialize();
  SecretKey sk2;
  sk2.Deserialize(skb);
  sk2.Init(pk2);
  EXPECT_TRUE(sk2 == k.sk);
  EXPECT_TRUE(sk2.lambda_ == k.sk.lambda_ && sk2.mu_ == k.sk.mu_);
  EXPECT_EQ(skb[0], 0x94);  // fixarray(4)
  Encryptor enc(k.pk);
  Decryptor dec2(pk2, sk2);
  std::vector<uint8_t> last;
  for (int64_t m : {(int64_t)0, (int64_t)1, (int64_t)-1, (int64_t)123456789, std::numeric_limits<int64_t>::min(),
                    std::numeric_limits<int64_t>::max()}) {
    const Ciphertext c = enc.Encrypt(BigInt(m));
    const auto b = c.Serialize(k.pk);
    EXPECT_EQ(b[0], 0x91);  // fixarray(1): MSGPACK_DEFINE(c_)
    EXPECT_TRUE(str_value(b, 1) == mod_mul(ct_value(c), R, n2));
    const Ciphertext c2 = Ciphertext::Deserialize(pk2, b);
    EXPECT_TRUE(c2 == c);
    EXPECT_TRUE(dec2.Decrypt(c2) == BigInt(m));
    last = b;
  }
  last.pop_back();
  EXPECT_THROW(Ciphertext::Deserialize(pk2, last));
  // phe/phe_test/ser_test.cc:31-108: keys and ciphertexts print with content,
  // and a party holding only the public key's bytes evaluates on a client's
  // ciphertext (AddInplace, Randomize) and sends it back
  EXPECT_TRUE(k.pk.ToString().size() > 30 && k.sk.ToString().size() > 30);
  const Ciphertext client = enc.Encrypt(BigInt(-963258741));
  EXPECT_TRUE(client.ToString().size() >= 10);
  PublicKey server_pk;
  server_pk.Deserialize(k.pk.Serialize());
  Evaluator server(server_pk);
  Ciphertext at_server = Ciphertext::Deserialize(server_pk, client.Serialize(k.pk));
  server.AddInplace(&at_server, BigInt(666));
  server.Randomize(&at_server);
  const Ciphertext back = Ciphertext::Deserialize(k.pk, at_server.Serialize(server_pk));
  EXPECT_TRUE(dec2.Decrypt(back) == BigInt(-963258741 + 666));
  const auto audit = enc.EncryptWithAudit(BigInt(-77));
  const std::string want_c = "c:" + mod_mul(ct_value(audit.first), R, n2).ToHexString();
  EXPECT_TRUE(audit.second.rfind("p:-4D,rn:", 0) == 0);
  EXPECT_TRUE(audit.second.size() > want_c.size() &&
              audit.second.compare(audit.second.size() - want_c.size(), want_c.size(), want_c) == 0);
}

int main() {
  Keys k2048, k3072, seeded;
  run("KeyGenerator.Generate(2048)", [&] {
    KeyGenerator::Generate(2048, &k2048.sk, &k2048.pk);
    EXPECT_EQ(k2048.pk.n_.BitCount(), (size_t)2048);
    EXPECT_TRUE(k2048.sk.p_ * k2048.sk.q_ == k2048.pk.n_);
  });
  run("KeyGenerator.GenerateSeeded(2048, 0)", [&] {
    KeyGenerator::GenerateSeeded(2048, 0, &seeded.sk, &seeded.pk);
    Keys again;
    KeyGenerator::GenerateSeeded(2048, 0, &again.sk, &again.pk);
    EXPECT_TRUE(seeded.pk == again.pk);
    EXPECT_TRUE(seeded.sk == again.sk);
    std::printf("  seeded n = %s\n", seeded.pk.n_.ToHexString().c_str());
  });
  run("ZPaillierTest.CiphertextEvaluate", [&] { CiphertextEvaluate(k2048); });
  run("ZPaillierTest.MinMaxDecrypt", [&] { MinMaxDecrypt(k2048); });
  run("ZPaillierTest.PlaintextEvaluate", [&] { PlaintextEvaluate(k2048); });
  run("BigNumberTest.EncDec+SubTest", [&] { BigNumberEncDecSub(k2048); });
  run("Vectorized.ClosedForms(2048, 257)", [&] { VectorizedClosedForms(seeded, 257); });
  run("KeyGenerator.Generate(3072)", [&] { KeyGenerator::Generate(3072, &k3072.sk, &k3072.pk); });
  run("ZPaillierTest.MinMaxDecrypt(3072)", [&] { MinMaxDecrypt(k3072); });
  run("Vectorized.ClosedForms(3072, 65)", [&] { VectorizedClosedForms(k3072, 65); });
  run("GPUTest cases of paillier_gpu_test.cc (2048)", [&] { GpuBackendCases(k2048); });
  run("phe EvaluatorTest / EncryptorTest cases (2048)", [&] { PheCases(k2048); });
  run("Vectorized.SpansSplitOverThreads(2048, 50000)", [&] { SpansSplitOverThreads(seeded); });
  run("Concurrent chunks coalesce (2048 + 3072 keys)", [&] { ConcurrentChunks(seeded, k3072); });
  run("Vectorized.SpanRates(2048, 65536)", [&] { SpanRates(seeded, 65536); });
  run("Decryptor rejects an unpaired key", [&] { EXPECT_THROW(Decryptor(k2048.pk, k3072.sk)); });
  run("Serialization(2048)", [&] { Serialization(k2048); });
  run("Serialization(3072)", [&] { Serialization(k3072); });
  if (g_failures) {
    std::printf("%d FAILURE(S)\n", g_failures);
    return 1;
  }
  std::printf("ALL PASSED\n");
  return 0;
}
