// keygen.cpp -- key generation for the engine (host, OpenSSL BIGNUM).
//
// Follows heu/library/algorithms/paillier_zahlen/key_generator.cc:26-60:
// Blum primes p, q of key_bits/2 bits with gcd(p-1, q-1) = 2 and
// bits(p - q) >= key_bits/2 - 2, bits(n) >= key_bits, x coprime to n,
// h = -x^2 mod n, h_s = h^n mod n^2 computed by CRT (secret_key.cc:45-62).
//
// The reference draws from the OS RNG.  Here randomness comes either from the
// OS (RAND_bytes) or from a seeded counter stream so keys are reproducible:
//   word(seed, stream, i) = mix(mix(seed*G + mix(stream + C)) + (i+1)*G)
// with mix = the SplitMix64 finaliser, G = 0x9E3779B97F4A7C15 and
// C = 0x632BE59BD9B4E019.  Prime search: a random start with the top bit and
// the two low bits set, then start, start+4, ... until a probable prime
// (trial division by the primes below 256, then Miller-Rabin to the first 24
// prime bases); stream 1 feeds p and q, stream 2 feeds x.
#include <openssl/bn.h>
#include <openssl/rand.h>

#include <cstdint>
#include <vector>

#include "../../include/heu_paillier.h"

namespace {

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct Stream {
  bool seeded;
  uint64_t key, ctr = 0;
  Stream(const uint64_t* seed, uint64_t stream)
      : seeded(seed != nullptr),
        key(seed ? mix64(*seed * 0x9E3779B97F4A7C15ULL + mix64(stream + 0x632BE59BD9B4E019ULL)) : 0) {}
  uint64_t next() {
    if (!seeded) {
      uint64_t v = 0;
      RAND_bytes(reinterpret_cast<unsigned char*>(&v), sizeof(v));
      return v;
    }
    return mix64(key + (++ctr) * 0x9E3779B97F4A7C15ULL);
  }
};

const unsigned kSmall[] = {3,   5,   7,   11,  13,  17,  19,  23,  29,  31,  37,  41,  43,  47,  53,  59,  61,  67,  71,
                           73,  79,  83,  89,  97,  101, 103, 107, 109, 113, 127, 131, 137, 139, 149, 151, 157, 163, 167,
                           173, 179, 181, 191, 193, 197, 199, 211, 223, 227, 229, 233, 239, 241, 251};

bool probable_prime(const BIGNUM* p, BN_CTX* ctx) {
  if (!BN_is_odd(p)) return BN_is_word(p, 2);
  for (unsigned s : kSmall) {
    if (BN_is_word(p, s)) return true;
    if (BN_mod_word(p, s) == 0) return false;
  }
  BN_CTX_start(ctx);
  BIGNUM *d = BN_CTX_get(ctx), *pm1 = BN_CTX_get(ctx), *a = BN_CTX_get(ctx), *x = BN_CTX_get(ctx);
  BN_copy(pm1, p);
  BN_sub_word(pm1, 1);
  BN_copy(d, pm1);
  int s = 0;
  while (!BN_is_odd(d)) {
    BN_rshift1(d, d);
    ++s;
  }
  bool prime = true;
  for (int i = 0; i < 24 && prime; ++i) {
    BN_set_word(a, i == 0 ? 2u : kSmall[i - 1]);
    BN_mod_exp(x, a, d, p, ctx);
    if (BN_is_one(x) || BN_cmp(x, pm1) == 0) continue;
    bool witness = true;
    for (int r = 1; r < s; ++r) {
      BN_mod_sqr(x, x, p, ctx);
      if (BN_cmp(x, pm1) == 0) {
        witness = false;
        break;
      }
    }
    if (witness) prime = false;
  }
  BN_CTX_end(ctx);
  return prime;
}

void random_bits(BIGNUM* out, int bits, Stream& st) {
  int words = (bits + 63) / 64;
  std::vector<uint64_t> buf(words);
  for (auto& w : buf) w = st.next();
  int top = bits - 64 * (words - 1);
  if (top < 64) buf[words - 1] &= (1ULL << top) - 1;
  buf[words - 1] |= 1ULL << (top - 1);
  BN_lebin2bn(reinterpret_cast<const unsigned char*>(buf.data()), words * 8, out);
}

void blum_prime(BIGNUM* p, int bits, Stream& st, BN_CTX* ctx) {
  for (;;) {
    random_bits(p, bits, st);
    BN_set_bit(p, 0);
    BN_set_bit(p, 1);
    for (int step = 0; step < 100000 && BN_num_bits(p) == bits; ++step) {
      if (probable_prime(p, ctx)) return;
      BN_add_word(p, 4);
    }
  }
}

void to_words(const BIGNUM* a, uint32_t* out, int words) {
  BN_bn2lebinpad(a, reinterpret_cast<unsigned char*>(out), words * 4);
}

}  // namespace

extern "C" int heu_keygen(uint32_t key_bits, const uint64_t* seed, uint32_t* n_out, uint32_t* hs_out, uint32_t* p_out,
                          uint32_t* q_out) {
  if (key_bits % 64 != 0 || key_bits < 128 || !n_out || !hs_out || !p_out || !q_out) return HEU_EINVAL;
  BN_CTX* ctx = BN_CTX_new();
  BIGNUM *p = BN_new(), *q = BN_new(), *n = BN_new(), *c = BN_new(), *t1 = BN_new(), *t2 = BN_new();
  const int half = (int)key_bits / 2;
  Stream sp(seed, 1), sx(seed, 2);
  do {
    blum_prime(p, half, sp, ctx);
    for (;;) {
      blum_prime(q, half, sp, ctx);
      BN_copy(t1, p);
      BN_sub_word(t1, 1);
      BN_copy(t2, q);
      BN_sub_word(t2, 1);
      BN_gcd(c, t1, t2, ctx);
      BN_sub(t1, p, q);
      if (BN_is_word(c, 2) && BN_num_bits(t1) >= half - 2) break;
    }
    BN_mul(n, p, q, ctx);
  } while (BN_num_bits(n) < (int)key_bits);
  BIGNUM *x = BN_new(), *h = BN_new();
  do {
    random_bits(x, (int)key_bits + 64, sx);
    BN_nnmod(x, x, n, ctx);
    BN_gcd(c, x, n, ctx);
  } while (!BN_is_one(c));
  // h = -x^2 mod n; h_s = h^n mod n^2 via the CRT halves mod p^2, q^2
  BN_mod_sqr(h, x, n, ctx);
  BN_sub(h, n, h);
  BIGNUM *p2 = BN_new(), *q2 = BN_new(), *n2 = BN_new(), *e = BN_new(), *hp = BN_new(), *hq = BN_new(),
         *hs = BN_new(), *inv = BN_new();
  BN_sqr(p2, p, ctx);
  BN_sqr(q2, q, ctx);
  BN_mul(n2, p2, q2, ctx);
  BN_copy(t1, p);
  BN_sub_word(t1, 1);
  BN_mul(t1, t1, p, ctx);  // p(p-1)
  BN_nnmod(e, n, t1, ctx);
  BN_nnmod(t2, h, p2, ctx);
  BN_mod_exp(hp, t2, e, p2, ctx);
  BN_copy(t1, q);
  BN_sub_word(t1, 1);
  BN_mul(t1, t1, q, ctx);  // q(q-1)
  BN_nnmod(e, n, t1, ctx);
  BN_nnmod(t2, h, q2, ctx);
  BN_mod_exp(hq, t2, e, q2, ctx);
  BN_mod_inverse(inv, q2, p2, ctx);
  BN_mul(inv, inv, q2, ctx);  // (q^2)^-1 mod p^2 * q^2
  BN_sub(hs, hp, hq);
  BN_mul(hs, hs, inv, ctx);
  BN_add(hs, hs, hq);
  BN_nnmod(hs, hs, n2, ctx);
  to_words(n, n_out, (int)key_bits / 32);
  to_words(hs, hs_out, (int)key_bits / 16);
  to_words(p, p_out, (int)key_bits / 64);
  to_words(q, q_out, (int)key_bits / 64);
  BIGNUM* all[] = {p, q, n, c, t1, t2, x, h, p2, q2, n2, e, hp, hq, hs, inv};
  for (BIGNUM* b : all) BN_free(b);
  BN_CTX_free(ctx);
  return HEU_OK;
}
