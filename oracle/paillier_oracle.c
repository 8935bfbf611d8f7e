/*
 * paillier_oracle.c -- CPU restatement of HEU's paillier_zahlen (ZPaillier)
 * over OpenSSL 3 BIGNUM.
 *
 * TEST INFRASTRUCTURE ONLY.  It checks the engine's results: only tests/, the
 * smoke() check of __graft_entry__.py and bench.py's cpu_baseline leg may
 * load it.  The product library (heu_amd/) never links or calls it.
 *
 * Every function names the reference file:line it restates.  Paths are
 * relative to the heu repository root.
 *
 * Conventions shared with the product C-ABI:
 *   - big integers cross the boundary as little-endian arrays of uint32
 *     limbs (u32 limb i holds bits [32i, 32i+32));
 *   - a ciphertext is the Z-space value c = (1 + n*m) * h_s^r mod n^2, not
 *     the reference's Montgomery-form c*R mod n^2 (the R of yacl's backend is
 *     not pinned); orc_to_mont / orc_from_mont export the M-space form for an
 *     explicit word size;
 *   - signed plaintexts cross as two's complement over `words` u32 limbs.
 *
 * Randomness: the reference draws keys and r from the OS RNG
 * (key_generator.cc:26-60, encryptor.cc:27).  Here every random word comes
 * from orc_prng_word(seed, stream, index), a counter-based SplitMix64, so the
 * same (seed, stream) reproduces the same keys and inputs everywhere.
 */
#include <openssl/bn.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_ERANGE 2
#define ORC_ENOTINV 3
#define ORC_ENOMEM 4

/* ------------------------------------------------------------------ */
/* Counter-based PRNG (SplitMix64 finaliser over a keyed counter).     */
/* ------------------------------------------------------------------ */
static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

uint64_t orc_prng_word(uint64_t seed, uint64_t stream, uint64_t index) {
  uint64_t key = mix64(seed * 0x9E3779B97F4A7C15ULL + mix64(stream + 0x632BE59BD9B4E019ULL));
  return mix64(key + (index + 1) * 0x9E3779B97F4A7C15ULL);
}

void orc_prng_fill(uint64_t seed, uint64_t stream, uint64_t offset, uint64_t *out, size_t count) {
  for (size_t i = 0; i < count; ++i) out[i] = orc_prng_word(seed, stream, offset + i);
}

/* ------------------------------------------------------------------ */
/* Limb <-> BIGNUM helpers                                             */
/* ------------------------------------------------------------------ */
static BIGNUM *bn_from_limbs(const uint32_t *limbs, size_t words) {
  return BN_lebin2bn((const unsigned char *)limbs, (int)(words * 4), NULL);
}

static int bn_to_limbs(const BIGNUM *a, uint32_t *limbs, size_t words) {
  if (BN_is_negative(a)) return ORC_ERANGE;
  return BN_bn2lebinpad(a, (unsigned char *)limbs, (int)(words * 4)) < 0 ? ORC_ERANGE : ORC_OK;
}

/* two's complement limbs -> signed BIGNUM */
static BIGNUM *bn_from_twos(const uint32_t *limbs, size_t words) {
  if (!(limbs[words - 1] & 0x80000000u)) return bn_from_limbs(limbs, words);
  uint32_t *tmp = (uint32_t *)malloc(words * 4);
  uint64_t carry = 1;
  for (size_t i = 0; i < words; ++i) {
    uint64_t v = (uint64_t)(uint32_t)~limbs[i] + carry;
    tmp[i] = (uint32_t)v;
    carry = v >> 32;
  }
  BIGNUM *r = bn_from_limbs(tmp, words);
  free(tmp);
  BN_set_negative(r, 1);
  return r;
}

/* signed BIGNUM -> two's complement limbs; ERANGE if it does not fit */
static int bn_to_twos(const BIGNUM *a, uint32_t *limbs, size_t words) {
  size_t bits = (size_t)BN_num_bits(a), top = words * 32 - 1;
  /* fits iff -2^top <= a < 2^top */
  if (bits > top + 1) return ORC_ERANGE;
  if (bits == top + 1) {
    if (!BN_is_negative(a)) return ORC_ERANGE;
    for (size_t b = 0; b < top; ++b)
      if (BN_is_bit_set(a, (int)b)) return ORC_ERANGE;
  }
  if (BN_bn2lebinpad(a, (unsigned char *)limbs, (int)(words * 4)) < 0) return ORC_ERANGE;
  if (BN_is_negative(a)) {
    uint64_t carry = 1;
    for (size_t i = 0; i < words; ++i) {
      uint64_t v = (uint64_t)(uint32_t)~limbs[i] + carry;
      limbs[i] = (uint32_t)v;
      carry = v >> 32;
    }
  }
  return ORC_OK;
}

static BIGNUM *bn_from_i64(int64_t v) {
  BIGNUM *r = BN_new();
  uint64_t mag = v < 0 ? (uint64_t)0 - (uint64_t)v : (uint64_t)v;
  BN_set_word(r, mag);
  if (v < 0) BN_set_negative(r, 1);
  return r;
}

/* non-negative residue of a mod m */
static void bn_nnmod(BIGNUM *r, const BIGNUM *a, const BIGNUM *m, BN_CTX *ctx) { BN_nnmod(r, a, m, ctx); }

/* ------------------------------------------------------------------ */
/* Primality: trial division + Miller-Rabin over fixed bases, so that  */
/* prime search is a deterministic function of the PRNG stream.        */
/* ------------------------------------------------------------------ */
static const unsigned kSmallPrimes[] = {3,   5,   7,   11,  13,  17,  19,  23,  29,  31,  37,  41,  43,  47,
                                        53,  59,  61,  67,  71,  73,  79,  83,  89,  97,  101, 103, 107, 109,
                                        113, 127, 131, 137, 139, 149, 151, 157, 163, 167, 173, 179, 181, 191,
                                        193, 197, 199, 211, 223, 227, 229, 233, 239, 241, 251};
#define ORC_MR_BASES 24

int orc_is_probable_prime(const BIGNUM *p, BN_CTX *ctx) {
  if (BN_is_word(p, 2)) return 1;
  if (!BN_is_odd(p) || BN_is_one(p) || BN_is_zero(p)) return 0;
  for (size_t i = 0; i < sizeof(kSmallPrimes) / sizeof(kSmallPrimes[0]); ++i) {
    if (BN_is_word(p, kSmallPrimes[i])) return 1;
    if (BN_mod_word(p, kSmallPrimes[i]) == 0) return 0;
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
  int prime = 1;
  unsigned bases[ORC_MR_BASES];
  bases[0] = 2;
  for (int i = 1; i < ORC_MR_BASES; ++i) bases[i] = kSmallPrimes[i - 1];
  for (int i = 0; i < ORC_MR_BASES && prime; ++i) {
    BN_set_word(a, bases[i]);
    BN_mod_exp(x, a, d, p, ctx);
    if (BN_is_one(x) || BN_cmp(x, pm1) == 0) continue;
    int composite = 1;
    for (int r = 1; r < s; ++r) {
      BN_mod_sqr(x, x, p, ctx);
      if (BN_cmp(x, pm1) == 0) {
        composite = 0;
        break;
      }
    }
    if (composite) prime = 0;
  }
  BN_CTX_end(ctx);
  return prime;
}

/* Random `bits`-bit number from the stream: little-endian words, top bit set. */
static void rand_bits(BIGNUM *out, int bits, uint64_t seed, uint64_t stream, uint64_t *ctr) {
  int words = (bits + 63) / 64;
  uint64_t *buf = (uint64_t *)calloc((size_t)words, 8);
  for (int i = 0; i < words; ++i) buf[i] = orc_prng_word(seed, stream, (*ctr)++);
  int top = bits - 64 * (words - 1); /* bits used in the last word, 1..64 */
  if (top < 64) buf[words - 1] &= (1ULL << top) - 1;
  buf[words - 1] |= 1ULL << (top - 1);
  BN_lebin2bn((const unsigned char *)buf, words * 8, out);
  free(buf);
}

/* Blum prime (p = 3 mod 4) with exactly `bits` bits.  Mirrors the role of
 * yacl BigInt::RandPrimeOver(bits, PrimeType::BBS) used at
 * paillier_zahlen/key_generator.cc:33,35: a random start from the stream,
 * then the first probable prime in start, start+4, start+8, ... */
static void rand_bbs_prime(BIGNUM *p, int bits, uint64_t seed, uint64_t stream, uint64_t *ctr, BN_CTX *ctx) {
  for (;;) {
    rand_bits(p, bits, seed, stream, ctr);
    /* force p = 3 mod 4 */
    BN_set_bit(p, 0);
    BN_set_bit(p, 1);
    for (int step = 0; step < 100000; ++step) {
      if (BN_num_bits(p) != bits) break;
      if (orc_is_probable_prime(p, ctx)) return;
      BN_add_word(p, 4);
    }
  }
}

/* ------------------------------------------------------------------ */
/* Key object: PublicKey::Init (public_key.cc:28-40) + SecretKey::Init */
/* (secret_key.cc:19-43) derived values.                               */
/* ------------------------------------------------------------------ */
typedef struct orc_key {
  int key_bits;          /* bits(n) */
  int n_words, ct_words; /* u32 limbs of n and n^2 */
  int has_sk;
  BIGNUM *n, *n2, *n_half, *hs;
  BN_MONT_CTX *mont_n2;
  /* fixed-base table of h_s (public_key.cc:36-39): table[w][d] = hs^(d*2^(unit*w)) in M-space */
  int unit_bits, table_windows;
  BIGNUM **table;
  /* secret key */
  BIGNUM *p, *q, *p2, *q2, *lambda, *mu, *q2_inv_mul_q2, *p_inv_mod_q, *phi_p2, *phi_q2, *phi_p, *phi_q, *hp, *hq;
  BN_MONT_CTX *mont_p2, *mont_q2;
} orc_key;

/* secret_key.cc:45-62 PowModNSquareCrt: base^exp mod n^2 via the halves mod p^2, q^2 */
static void pow_mod_nsquare_crt(BIGNUM *out, const orc_key *k, const BIGNUM *base, const BIGNUM *exp, BN_CTX *ctx) {
  BN_CTX_start(ctx);
  BIGNUM *pexp = BN_CTX_get(ctx), *qexp = BN_CTX_get(ctx), *pb = BN_CTX_get(ctx), *qb = BN_CTX_get(ctx);
  BIGNUM *pr = BN_CTX_get(ctx), *qr = BN_CTX_get(ctx), *t = BN_CTX_get(ctx);
  bn_nnmod(pexp, exp, k->phi_p2, ctx);
  bn_nnmod(qexp, exp, k->phi_q2, ctx);
  bn_nnmod(pb, base, k->p2, ctx);
  bn_nnmod(qb, base, k->q2, ctx);
  BN_mod_exp(pr, pb, pexp, k->p2, ctx);
  BN_mod_exp(qr, qb, qexp, k->q2, ctx);
  BN_sub(t, pr, qr);
  BN_mul(t, t, k->q2_inv_mul_q2, ctx);
  BN_add(t, t, qr);
  bn_nnmod(out, t, k->n2, ctx);
  BN_CTX_end(ctx);
}

static void secret_key_init(orc_key *k, BN_CTX *ctx) {
  /* secret_key.cc:19-43 */
  k->p2 = BN_new();
  k->q2 = BN_new();
  BN_sqr(k->p2, k->p, ctx);
  BN_sqr(k->q2, k->q, ctx);
  BIGNUM *q2inv = BN_mod_inverse(NULL, k->q2, k->p2, ctx);
  k->q2_inv_mul_q2 = BN_new();
  BN_mul(k->q2_inv_mul_q2, q2inv, k->q2, ctx);
  BN_free(q2inv);
  k->p_inv_mod_q = BN_mod_inverse(NULL, k->p, k->q, ctx);
  k->phi_p = BN_dup(k->p);
  BN_sub_word(k->phi_p, 1);
  k->phi_q = BN_dup(k->q);
  BN_sub_word(k->phi_q, 1);
  k->phi_p2 = BN_new();
  BN_mul(k->phi_p2, k->p, k->phi_p, ctx);
  k->phi_q2 = BN_new();
  BN_mul(k->phi_q2, k->q, k->phi_q, ctx);
  /* hp = (L_p((n+1)^(p-1) mod p^2))^-1 mod p, likewise hq */
  BIGNUM *g = BN_new(), *t = BN_new(), *rem = BN_new();
  BN_mul(g, k->p, k->q, ctx);
  BN_add_word(g, 1);
  BN_mod_exp(t, g, k->phi_p, k->p2, ctx);
  BN_sub_word(t, 1);
  BN_div(t, rem, t, k->p, ctx);
  k->hp = BN_mod_inverse(NULL, t, k->p, ctx);
  BN_mod_exp(t, g, k->phi_q, k->q2, ctx);
  BN_sub_word(t, 1);
  BN_div(t, rem, t, k->q, ctx);
  k->hq = BN_mod_inverse(NULL, t, k->q, ctx);
  BN_free(g);
  BN_free(t);
  BN_free(rem);
  /* lambda = (p-1)(q-1)/2, mu = lambda^-1 mod n (key_generator.cc:51-52) */
  k->lambda = BN_new();
  BN_mul(k->lambda, k->phi_p, k->phi_q, ctx);
  BN_rshift1(k->lambda, k->lambda);
  k->mu = BN_mod_inverse(NULL, k->lambda, k->n, ctx);
  k->mont_p2 = BN_MONT_CTX_new();
  BN_MONT_CTX_set(k->mont_p2, k->p2, ctx);
  k->mont_q2 = BN_MONT_CTX_new();
  BN_MONT_CTX_set(k->mont_q2, k->q2, ctx);
  k->has_sk = 1;
}

static void public_key_init(orc_key *k, BN_CTX *ctx) {
  /* public_key.cc:28-34 */
  k->n2 = BN_new();
  BN_sqr(k->n2, k->n, ctx);
  k->n_half = BN_new();
  BN_rshift1(k->n_half, k->n);
  k->key_bits = BN_num_bits(k->n);
  k->n_words = (k->key_bits + 31) / 32;
  k->ct_words = 2 * k->n_words;
  k->mont_n2 = BN_MONT_CTX_new();
  BN_MONT_CTX_set(k->mont_n2, k->n2, ctx);
}

/* public_key.cc:36-39 MakeBaseTable(h_s, unit, max_exp_bits): table[w][d] =
 * h_s^(d * 2^(unit*w)) kept in Montgomery form for MulMod. */
int orc_key_build_table(orc_key *k, int unit_bits, int max_exp_bits) {
  if (unit_bits < 1 || unit_bits > 16 || max_exp_bits < 1) return ORC_EINVAL;
  BN_CTX *ctx = BN_CTX_new();
  int windows = (max_exp_bits + unit_bits - 1) / unit_bits;
  size_t per = (size_t)1 << unit_bits;
  BIGNUM **t = (BIGNUM **)calloc((size_t)windows * per, sizeof(BIGNUM *));
  BIGNUM *base = BN_new();
  BN_to_montgomery(base, k->hs, k->mont_n2, ctx);
  for (int w = 0; w < windows; ++w) {
    t[w * per + 0] = BN_new();
    BN_to_montgomery(t[w * per + 0], BN_value_one(), k->mont_n2, ctx);
    t[w * per + 1] = BN_dup(base);
    for (size_t d = 2; d < per; ++d) {
      t[w * per + d] = BN_new();
      BN_mod_mul_montgomery(t[w * per + d], t[w * per + d - 1], base, k->mont_n2, ctx);
    }
    /* next window's base = base^(2^unit) */
    BN_mod_mul_montgomery(base, t[w * per + per - 1], base, k->mont_n2, ctx);
  }
  BN_free(base);
  k->table = t;
  k->unit_bits = unit_bits;
  k->table_windows = windows;
  BN_CTX_free(ctx);
  return ORC_OK;
}

static orc_key *key_alloc(void) { return (orc_key *)calloc(1, sizeof(orc_key)); }

/* Public + (optional) secret key from limbs.  p, q may be NULL. */
orc_key *orc_key_new(const uint32_t *n, const uint32_t *hs, int n_words, const uint32_t *p, const uint32_t *q, int prime_words) {
  BN_CTX *ctx = BN_CTX_new();
  orc_key *k = key_alloc();
  k->n = bn_from_limbs(n, (size_t)n_words);
  public_key_init(k, ctx);
  k->hs = bn_from_limbs(hs, (size_t)k->ct_words);
  if (p && q) {
    k->p = bn_from_limbs(p, (size_t)prime_words);
    k->q = bn_from_limbs(q, (size_t)prime_words);
    secret_key_init(k, ctx);
  }
  BN_CTX_free(ctx);
  return k;
}

void orc_key_free(orc_key *k) {
  if (!k) return;
  BIGNUM *all[] = {k->n, k->n2, k->n_half, k->hs, k->p, k->q, k->p2, k->q2, k->lambda, k->mu, k->q2_inv_mul_q2,
                   k->p_inv_mod_q, k->phi_p2, k->phi_q2, k->phi_p, k->phi_q, k->hp, k->hq};
  for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); ++i) BN_free(all[i]);
  if (k->table) {
    size_t total = (size_t)k->table_windows << k->unit_bits;
    for (size_t i = 0; i < total; ++i) BN_free(k->table[i]);
    free(k->table);
  }
  BN_MONT_CTX_free(k->mont_n2);
  BN_MONT_CTX_free(k->mont_p2);
  BN_MONT_CTX_free(k->mont_q2);
  free(k);
}

int orc_key_bits(const orc_key *k) { return k->key_bits; }

/* paillier_zahlen/key_generator.cc:26-60 with the OS RNG replaced by the
 * seeded stream.  Outputs p, q (prime_words limbs each, prime_words =
 * key_bits/64), n (key_bits/32 limbs) and h_s (key_bits/16 limbs). */
int orc_keygen(int key_bits, uint64_t seed, uint32_t *p_out, uint32_t *q_out, uint32_t *n_out, uint32_t *hs_out) {
  if (key_bits % 64 != 0 || key_bits < 128) return ORC_EINVAL;
  BN_CTX *ctx = BN_CTX_new();
  BIGNUM *p = BN_new(), *q = BN_new(), *n = BN_new(), *c = BN_new(), *t1 = BN_new(), *t2 = BN_new();
  uint64_t ctr = 0;
  int half = key_bits / 2;
  do {
    rand_bbs_prime(p, half, seed, 1, &ctr, ctx);
    for (;;) {
      rand_bbs_prime(q, half, seed, 1, &ctr, ctx);
      BN_copy(t1, p);
      BN_sub_word(t1, 1);
      BN_copy(t2, q);
      BN_sub_word(t2, 1);
      BN_gcd(c, t1, t2, ctx);
      BN_sub(t1, p, q);
      /* kPQDifferenceBitLenSub = 2 (key_generator.cc:24) */
      if (BN_is_word(c, 2) && BN_num_bits(t1) >= half - 2) break;
    }
    BN_mul(n, p, q, ctx);
  } while (BN_num_bits(n) < key_bits);
  /* x coprime to n, h = -x^2 mod n (key_generator.cc:41-47) */
  BIGNUM *x = BN_new(), *h = BN_new();
  uint64_t xctr = 0;
  do {
    rand_bits(x, key_bits + 64, seed, 2, &xctr); /* uniform enough after reduction */
    BN_nnmod(x, x, n, ctx);
    BN_gcd(c, x, n, ctx);
  } while (!BN_is_one(c));
  BN_mod_sqr(h, x, n, ctx);
  BN_set_negative(h, !BN_is_zero(h));
  orc_key *k = key_alloc();
  k->n = BN_dup(n);
  k->p = BN_dup(p);
  k->q = BN_dup(q);
  public_key_init(k, ctx);
  secret_key_init(k, ctx);
  /* h_s = PowModNSquareCrt(h, n) (key_generator.cc:56) */
  k->hs = BN_new();
  pow_mod_nsquare_crt(k->hs, k, h, n, ctx);
  int rc = bn_to_limbs(p, p_out, (size_t)key_bits / 64) | bn_to_limbs(q, q_out, (size_t)key_bits / 64) |
           bn_to_limbs(n, n_out, (size_t)key_bits / 32) | bn_to_limbs(k->hs, hs_out, (size_t)key_bits / 16);
  orc_key_free(k);
  BN_free(p);
  BN_free(q);
  BN_free(n);
  BN_free(c);
  BN_free(t1);
  BN_free(t2);
  BN_free(x);
  BN_free(h);
  BN_CTX_free(ctx);
  return rc ? ORC_ERANGE : ORC_OK;
}

/* Export sk-derived values the engine needs, for cross-checks in tests. */
int orc_key_export(const orc_key *k, const char *name, uint32_t *out, int words) {
  const BIGNUM *v = NULL;
  if (!strcmp(name, "n2")) v = k->n2;
  else if (!strcmp(name, "n_half")) v = k->n_half;
  else if (!strcmp(name, "lambda")) v = k->lambda;
  else if (!strcmp(name, "mu")) v = k->mu;
  else if (!strcmp(name, "hp")) v = k->hp;
  else if (!strcmp(name, "hq")) v = k->hq;
  else if (!strcmp(name, "p_inv_mod_q")) v = k->p_inv_mod_q;
  else if (!strcmp(name, "q2_inv_mul_q2")) v = k->q2_inv_mul_q2;
  if (!v) return ORC_EINVAL;
  return bn_to_limbs(v, out, (size_t)words);
}

/* ------------------------------------------------------------------ */
/* Per-element operations                                             */
/* ------------------------------------------------------------------ */
typedef struct {
  const orc_key *k;
  BN_CTX *ctx;
} orc_tls;

/* encryptor.cc:26-31 GetRn: h_s^r, either through the fixed-base table in
 * Montgomery space (the reference's route) or by BN_mod_exp.  Output is the
 * Z-space value. */
static int get_rn(BIGNUM *out, const orc_key *k, const BIGNUM *r, int use_table, BN_CTX *ctx) {
  if (!use_table || !k->table) {
    BN_mod_exp(out, k->hs, r, k->n2, ctx);
    return ORC_OK;
  }
  int bits = BN_num_bits(r);
  if (bits > k->table_windows * k->unit_bits) return ORC_ERANGE;
  size_t per = (size_t)1 << k->unit_bits;
  int started = 0;
  for (int w = 0; w * k->unit_bits < bits; ++w) {
    unsigned d = 0;
    for (int b = k->unit_bits - 1; b >= 0; --b) d = (d << 1) | (unsigned)BN_is_bit_set(r, w * k->unit_bits + b);
    if (d == 0) continue;
    if (!started) {
      BN_copy(out, k->table[w * per + d]);
      started = 1;
    } else {
      BN_mod_mul_montgomery(out, out, k->table[w * per + d], k->mont_n2, ctx);
    }
  }
  if (!started) BN_to_montgomery(out, BN_value_one(), k->mont_n2, ctx);
  BN_from_montgomery(out, out, k->mont_n2, ctx);
  return ORC_OK;
}

/* encryptor.cc:35-49 EncryptImpl: |m| <= n/2 else error; gm = n*m + 1 (no
 * reduction; negative m gives (1+nm) mod n^2); c = gm * h_s^r mod n^2. */
static int encrypt_one(BIGNUM *ct, const orc_key *k, const BIGNUM *m, const BIGNUM *r, int use_table, BN_CTX *ctx) {
  BN_CTX_start(ctx);
  BIGNUM *absm = BN_CTX_get(ctx), *gm = BN_CTX_get(ctx), *rn = BN_CTX_get(ctx);
  BN_copy(absm, m);
  BN_set_negative(absm, 0);
  int rc = ORC_OK;
  if (BN_cmp(absm, k->n_half) > 0) {
    rc = ORC_ERANGE;
    goto done;
  }
  BN_mul(gm, k->n, m, ctx);
  BN_add_word(gm, 1);
  bn_nnmod(gm, gm, k->n2, ctx);
  rc = get_rn(rn, k, r, use_table, ctx);
  if (rc) goto done;
  BN_mod_mul(ct, gm, rn, k->n2, ctx);
done:
  BN_CTX_end(ctx);
  return rc;
}

/* decryptor.cc:32-53: CRT decryption with sign fold at n/2. */
static int decrypt_one(BIGNUM *out, const orc_key *k, const BIGNUM *c, BN_CTX *ctx) {
  if (BN_is_negative(c) || BN_cmp(c, k->n2) >= 0) return ORC_ERANGE;
  BN_CTX_start(ctx);
  BIGNUM *mp = BN_CTX_get(ctx), *mq = BN_CTX_get(ctx), *rem = BN_CTX_get(ctx), *t = BN_CTX_get(ctx);
  BN_mod_exp_mont(mp, c, k->phi_p, k->p2, ctx, k->mont_p2);
  BN_sub_word(mp, 1);
  BN_div(mp, rem, mp, k->p, ctx);
  BN_mod_mul(mp, mp, k->hp, k->p, ctx);
  BN_mod_exp_mont(mq, c, k->phi_q, k->q2, ctx, k->mont_q2);
  BN_sub_word(mq, 1);
  BN_div(mq, rem, mq, k->q, ctx);
  BN_mod_mul(mq, mq, k->hq, k->q, ctx);
  BN_sub(t, mq, mp);
  BN_nnmod(t, t, k->q, ctx); /* MulMod reduces into [0, q) */
  BN_mod_mul(t, t, k->p_inv_mod_q, k->q, ctx);
  BN_mul(t, t, k->p, ctx);
  BN_add(out, t, mp);
  if (BN_cmp(out, k->n_half) > 0) BN_sub(out, out, k->n);
  BN_CTX_end(ctx);
  return ORC_OK;
}

/* evaluator.cc:82-91 Negate: c^-1 mod n^2 */
static int negate_one(BIGNUM *out, const orc_key *k, const BIGNUM *c, BN_CTX *ctx) {
  BIGNUM *r = BN_mod_inverse(out, c, k->n2, ctx);
  return r ? ORC_OK : ORC_ENOTINV;
}

/* evaluator.cc:99-123 Mul: bits(p)=0 -> Identity (1); p=1 -> a; p=-1 ->
 * Negate(a); otherwise c^p mod n^2 with a negative p as an inverse power. */
static int mul_one(BIGNUM *out, const orc_key *k, const BIGNUM *c, const BIGNUM *p, BN_CTX *ctx) {
  int bits = BN_num_bits(p);
  if (bits == 0) {
    BN_one(out);
    return ORC_OK;
  }
  if (bits == 1) {
    if (BN_is_negative(p)) return negate_one(out, k, c, ctx);
    BN_copy(out, c);
    return ORC_OK;
  }
  BN_CTX_start(ctx);
  BIGNUM *e = BN_CTX_get(ctx);
  BN_copy(e, p);
  BN_set_negative(e, 0);
  BN_mod_exp(out, c, e, k->n2, ctx);
  int rc = ORC_OK;
  if (BN_is_negative(p)) rc = negate_one(out, k, out, ctx);
  BN_CTX_end(ctx);
  return rc;
}

/* ------------------------------------------------------------------ */
/* Batch drivers: one contiguous chunk per thread, as                  */
/* heu/library/numpy/encryptor.cc:42-52 splits work across workers.    */
/* ------------------------------------------------------------------ */
typedef struct job job;
typedef int (*elem_fn)(job *, size_t i, BN_CTX *ctx);
struct job {
  const orc_key *k;
  elem_fn fn;
  size_t begin, end;
  int rc;
  /* operands */
  const void *a, *b, *c;
  void *out;
  int aw, bw, ow, use_table;
};

static void *job_run(void *arg) {
  job *j = (job *)arg;
  BN_CTX *ctx = BN_CTX_new();
  for (size_t i = j->begin; i < j->end && !j->rc; ++i) j->rc = j->fn(j, i, ctx);
  BN_CTX_free(ctx);
  return NULL;
}

static int run_batch(job proto, size_t count, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  if ((size_t)nthreads > count) nthreads = count ? (int)count : 1;
  job *jobs = (job *)calloc((size_t)nthreads, sizeof(job));
  pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  size_t chunk = (count + (size_t)nthreads - 1) / (size_t)nthreads;
  for (int t = 0; t < nthreads; ++t) {
    jobs[t] = proto;
    jobs[t].begin = (size_t)t * chunk < count ? (size_t)t * chunk : count;
    jobs[t].end = jobs[t].begin + chunk < count ? jobs[t].begin + chunk : count;
    jobs[t].rc = 0;
    if (nthreads > 1) pthread_create(&th[t], NULL, job_run, &jobs[t]);
    else job_run(&jobs[t]);
  }
  int rc = 0;
  for (int t = 0; t < nthreads; ++t) {
    if (nthreads > 1) pthread_join(th[t], NULL);
    if (jobs[t].rc && !rc) rc = jobs[t].rc;
  }
  free(jobs);
  free(th);
  return rc;
}

static int el_encrypt_i64(job *j, size_t i, BN_CTX *ctx) {
  const orc_key *k = j->k;
  BIGNUM *m = bn_from_i64(((const int64_t *)j->a)[i]);
  BIGNUM *r = bn_from_limbs((const uint32_t *)j->b + i * (size_t)j->bw, (size_t)j->bw);
  BIGNUM *c = BN_new();
  int rc = encrypt_one(c, k, m, r, j->use_table, ctx);
  if (!rc) rc = bn_to_limbs(c, (uint32_t *)j->out + i * (size_t)k->ct_words, (size_t)k->ct_words);
  BN_free(m);
  BN_free(r);
  BN_free(c);
  return rc;
}

static int el_encrypt_twos(job *j, size_t i, BN_CTX *ctx) {
  const orc_key *k = j->k;
  BIGNUM *m = bn_from_twos((const uint32_t *)j->a + i * (size_t)j->aw, (size_t)j->aw);
  BIGNUM *r = bn_from_limbs((const uint32_t *)j->b + i * (size_t)j->bw, (size_t)j->bw);
  BIGNUM *c = BN_new();
  int rc = encrypt_one(c, k, m, r, j->use_table, ctx);
  if (!rc) rc = bn_to_limbs(c, (uint32_t *)j->out + i * (size_t)k->ct_words, (size_t)k->ct_words);
  BN_free(m);
  BN_free(r);
  BN_free(c);
  return rc;
}

static int el_decrypt(job *j, size_t i, BN_CTX *ctx) {
  const orc_key *k = j->k;
  BIGNUM *c = bn_from_limbs((const uint32_t *)j->a + i * (size_t)k->ct_words, (size_t)k->ct_words);
  BIGNUM *m = BN_new();
  int rc = decrypt_one(m, k, c, ctx);
  if (!rc) rc = bn_to_twos(m, (uint32_t *)j->out + i * (size_t)j->ow, (size_t)j->ow);
  BN_free(c);
  BN_free(m);
  return rc;
}

static int el_add(job *j, size_t i, BN_CTX *ctx) {
  /* evaluator.cc:30-37 Add(ct, ct) = MulMod */
  const orc_key *k = j->k;
  size_t w = (size_t)k->ct_words;
  BIGNUM *a = bn_from_limbs((const uint32_t *)j->a + i * w, w), *b = bn_from_limbs((const uint32_t *)j->b + i * w, w);
  BN_mod_mul(a, a, b, k->n2, ctx);
  int rc = bn_to_limbs(a, (uint32_t *)j->out + i * w, w);
  BN_free(a);
  BN_free(b);
  return rc;
}

static int el_add_pt_i64(job *j, size_t i, BN_CTX *ctx) {
  /* evaluator.cc:46-64 Add(ct, pt) = c * (1 + n*p) mod n^2, |p| <= n/2 */
  const orc_key *k = j->k;
  size_t w = (size_t)k->ct_words;
  BIGNUM *a = bn_from_limbs((const uint32_t *)j->a + i * w, w);
  BIGNUM *p = bn_from_i64(((const int64_t *)j->b)[i]), *gm = BN_new();
  BN_mul(gm, k->n, p, ctx);
  BN_add_word(gm, 1);
  bn_nnmod(gm, gm, k->n2, ctx);
  BN_mod_mul(a, a, gm, k->n2, ctx);
  int rc = bn_to_limbs(a, (uint32_t *)j->out + i * w, w);
  BN_free(a);
  BN_free(p);
  BN_free(gm);
  return rc;
}

static int el_sub(job *j, size_t i, BN_CTX *ctx) {
  /* evaluator.cc:66-68 Sub(a, b) = Add(a, Negate(b)) */
  const orc_key *k = j->k;
  size_t w = (size_t)k->ct_words;
  BIGNUM *a = bn_from_limbs((const uint32_t *)j->a + i * w, w), *b = bn_from_limbs((const uint32_t *)j->b + i * w, w);
  int rc = negate_one(b, k, b, ctx);
  if (!rc) {
    BN_mod_mul(a, a, b, k->n2, ctx);
    rc = bn_to_limbs(a, (uint32_t *)j->out + i * w, w);
  }
  BN_free(a);
  BN_free(b);
  return rc;
}

static int el_negate(job *j, size_t i, BN_CTX *ctx) {
  const orc_key *k = j->k;
  size_t w = (size_t)k->ct_words;
  BIGNUM *a = bn_from_limbs((const uint32_t *)j->a + i * w, w);
  int rc = negate_one(a, k, a, ctx);
  if (!rc) rc = bn_to_limbs(a, (uint32_t *)j->out + i * w, w);
  BN_free(a);
  return rc;
}

static int el_mul_i64(job *j, size_t i, BN_CTX *ctx) {
  const orc_key *k = j->k;
  size_t w = (size_t)k->ct_words;
  BIGNUM *a = bn_from_limbs((const uint32_t *)j->a + i * w, w), *p = bn_from_i64(((const int64_t *)j->b)[i]);
  BIGNUM *o = BN_new();
  int rc = mul_one(o, k, a, p, ctx);
  if (!rc) rc = bn_to_limbs(o, (uint32_t *)j->out + i * w, w);
  BN_free(a);
  BN_free(p);
  BN_free(o);
  return rc;
}

static int el_mul_twos(job *j, size_t i, BN_CTX *ctx) {
  const orc_key *k = j->k;
  size_t w = (size_t)k->ct_words;
  BIGNUM *a = bn_from_limbs((const uint32_t *)j->a + i * w, w);
  BIGNUM *p = bn_from_twos((const uint32_t *)j->b + i * (size_t)j->bw, (size_t)j->bw);
  BIGNUM *o = BN_new();
  int rc = mul_one(o, k, a, p, ctx);
  if (!rc) rc = bn_to_limbs(o, (uint32_t *)j->out + i * w, w);
  BN_free(a);
  BN_free(p);
  BN_free(o);
  return rc;
}

static int el_randomize(job *j, size_t i, BN_CTX *ctx) {
  /* evaluator.cc:25-28 Randomize: c * h_s^r */
  const orc_key *k = j->k;
  size_t w = (size_t)k->ct_words;
  BIGNUM *a = bn_from_limbs((const uint32_t *)j->a + i * w, w);
  BIGNUM *r = bn_from_limbs((const uint32_t *)j->b + i * (size_t)j->bw, (size_t)j->bw), *rn = BN_new();
  int rc = get_rn(rn, k, r, j->use_table, ctx);
  if (!rc) {
    BN_mod_mul(a, a, rn, k->n2, ctx);
    rc = bn_to_limbs(a, (uint32_t *)j->out + i * w, w);
  }
  BN_free(a);
  BN_free(r);
  BN_free(rn);
  return rc;
}

int orc_encrypt_i64(const orc_key *k, const int64_t *m, const uint32_t *r, int r_words, size_t count, uint32_t *ct,
                    int use_table, int nthreads) {
  job p = {k, el_encrypt_i64, 0, 0, 0, m, r, NULL, ct, 0, r_words, 0, use_table};
  return run_batch(p, count, nthreads);
}

int orc_encrypt_twos(const orc_key *k, const uint32_t *m, int m_words, const uint32_t *r, int r_words, size_t count,
                     uint32_t *ct, int use_table, int nthreads) {
  job p = {k, el_encrypt_twos, 0, 0, 0, m, r, NULL, ct, m_words, r_words, 0, use_table};
  return run_batch(p, count, nthreads);
}

int orc_decrypt(const orc_key *k, const uint32_t *ct, size_t count, uint32_t *out, int out_words, int nthreads) {
  if (!k->has_sk) return ORC_EINVAL;
  job p = {k, el_decrypt, 0, 0, 0, ct, NULL, NULL, out, 0, 0, out_words, 0};
  return run_batch(p, count, nthreads);
}

int orc_add(const orc_key *k, const uint32_t *a, const uint32_t *b, size_t count, uint32_t *out, int nthreads) {
  job p = {k, el_add, 0, 0, 0, a, b, NULL, out, 0, 0, 0, 0};
  return run_batch(p, count, nthreads);
}

int orc_add_pt_i64(const orc_key *k, const uint32_t *a, const int64_t *pt, size_t count, uint32_t *out, int nthreads) {
  job p = {k, el_add_pt_i64, 0, 0, 0, a, pt, NULL, out, 0, 0, 0, 0};
  return run_batch(p, count, nthreads);
}

int orc_sub(const orc_key *k, const uint32_t *a, const uint32_t *b, size_t count, uint32_t *out, int nthreads) {
  job p = {k, el_sub, 0, 0, 0, a, b, NULL, out, 0, 0, 0, 0};
  return run_batch(p, count, nthreads);
}

int orc_negate(const orc_key *k, const uint32_t *a, size_t count, uint32_t *out, int nthreads) {
  job p = {k, el_negate, 0, 0, 0, a, NULL, NULL, out, 0, 0, 0, 0};
  return run_batch(p, count, nthreads);
}

int orc_mul_i64(const orc_key *k, const uint32_t *a, const int64_t *s, size_t count, uint32_t *out, int nthreads) {
  job p = {k, el_mul_i64, 0, 0, 0, a, s, NULL, out, 0, 0, 0, 0};
  return run_batch(p, count, nthreads);
}

/* Mul with scalars of any length, two's complement over s_words limbs. */
int orc_mul_twos(const orc_key *k, const uint32_t *a, const uint32_t *s, int s_words, size_t count, uint32_t *out,
                 int nthreads) {
  job p = {k, el_mul_twos, 0, 0, 0, a, s, NULL, out, 0, s_words, 0, 0};
  return run_batch(p, count, nthreads);
}

int orc_randomize(const orc_key *k, const uint32_t *a, const uint32_t *r, int r_words, size_t count, uint32_t *out,
                  int use_table, int nthreads) {
  job p = {k, el_randomize, 0, 0, 0, a, r, NULL, out, 0, r_words, 0, use_table};
  return run_batch(p, count, nthreads);
}

/* numpy/evaluator.cc:355-373 Sum: the product of the ciphertexts of each
 * segment [offsets[s], offsets[s+1]) mod n^2; an empty segment gives the
 * identity 1.  If weights is non-NULL each term is Mul(c_i, w_i) first
 * (numpy/evaluator.cc:271-292 DoCallMatMul scalar path: Mul then AddInplace). */
int orc_dot_i64(const orc_key *k, const uint32_t *cts, const int64_t *weights, const int64_t *offsets, size_t segments,
                uint32_t *out) {
  BN_CTX *ctx = BN_CTX_new();
  size_t w = (size_t)k->ct_words;
  BIGNUM *acc = BN_new(), *c = BN_new(), *t = BN_new(), *p = NULL;
  int rc = ORC_OK;
  for (size_t s = 0; s < segments && !rc; ++s) {
    BN_one(acc);
    for (int64_t i = offsets[s]; i < offsets[s + 1] && !rc; ++i) {
      BN_lebin2bn((const unsigned char *)(cts + (size_t)i * w), (int)(w * 4), c);
      if (weights) {
        BN_free(p);
        p = bn_from_i64(weights[i]);
        rc = mul_one(t, k, c, p, ctx);
        BN_mod_mul(acc, acc, t, k->n2, ctx);
      } else {
        BN_mod_mul(acc, acc, c, k->n2, ctx);
      }
    }
    if (!rc) rc = bn_to_limbs(acc, out + s * w, w);
  }
  BN_free(p);
  BN_free(acc);
  BN_free(c);
  BN_free(t);
  BN_CTX_free(ctx);
  return rc;
}

/* M-space export: c*R mod n^2 with R = 2^(word_bits * ceil(bits(n^2)/word_bits)),
 * the MontgomerySpace representation stored in paillier_z::Ciphertext::c_
 * (encryptor.cc:46, decryptor.cc:36). */
int orc_to_mont(const orc_key *k, const uint32_t *in, size_t count, int word_bits, int to_mont, uint32_t *out) {
  BN_CTX *ctx = BN_CTX_new();
  size_t w = (size_t)k->ct_words;
  int rbits = (BN_num_bits(k->n2) + word_bits - 1) / word_bits * word_bits;
  BIGNUM *R = BN_new(), *Rinv = NULL, *x = BN_new();
  BN_set_bit(R, rbits);
  BN_nnmod(R, R, k->n2, ctx);
  if (!to_mont) Rinv = BN_mod_inverse(NULL, R, k->n2, ctx);
  int rc = ORC_OK;
  for (size_t i = 0; i < count && !rc; ++i) {
    BN_lebin2bn((const unsigned char *)(in + i * w), (int)(w * 4), x);
    BN_mod_mul(x, x, to_mont ? R : Rinv, k->n2, ctx);
    rc = bn_to_limbs(x, out + i * w, w);
  }
  BN_free(R);
  BN_free(Rinv);
  BN_free(x);
  BN_CTX_free(ctx);
  return rc;
}
