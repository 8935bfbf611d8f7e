// paillier_device.hpp -- gfx950 kernels of the Paillier batch engine.
//
// Every kernel works on residues spread over lane groups (mp_group.hpp).  One
// group handles one element (or one chunk of a reduction); a block holds
// kBlock/L groups.  The modulus is staged in LDS once per block.  Each group
// owns two LDS operand slots: A holds the running value (the multiplier operand
// of every Montgomery product), B parks a second value; registers only carry
// the Montgomery working set, which is what lets the wide shapes (K = 27, 36)
// run without spilling.  Kernels never allocate: scratch buffers come from the
// host engine.
#pragma once
#include "kernels.hpp"
#include "mp_group.hpp"

namespace heu_hip {

constexpr int kPowWin = kPowWindowBits;
// Decryption window for a ciphertext of W words (key bits = 16 W).
template <int W>
constexpr int kDecWin = dec_window_bits(16 * W);

template <int L, int K, int W>
struct Cfg {
  static constexpr int NL = L * K;
  static constexpr int kHalfW = (W + 1) / 2;
  // slot words: >= NL, and two slots hold a staged W-word number
  static constexpr int S = (((NL > kHalfW ? NL : kHalfW) + 3) & ~3) + 4;
  static constexpr int kGroups = kBlock / L;
  static constexpr int kLdsWords = NL + kGroups * 2 * S;
  // registers of the Montgomery working set: 2K (accumulators) + K (b) + K (m)
  static constexpr int kRegWaves = 4 * K + 40 <= 128 ? 4 : (4 * K + 40 <= 168 ? 3 : 2);
  // a block is one wave per SIMD; LDS (160 KiB per CU) caps resident blocks,
  // and a register budget tighter than that cap only adds spills
  static constexpr int kLdsBlocks = (160 * 1024) / (kLdsWords * 4 + 64);
  static constexpr int kWaves = kRegWaves < kLdsBlocks ? kRegWaves : (kLdsBlocks < 1 ? 1 : kLdsBlocks);
};

template <int L, int K>
__device__ __forceinline__ void one_limbs(u32 (&x)[K], u32 lid) {
#pragma unroll
  for (int t = 0; t < K; ++t) x[t] = (lid == 0u && t == 0) ? 1u : 0u;
}

// Group context: the two LDS slots, the modulus and the lane.
template <int L, int K, int W>
struct Grp {
  typedef Cfg<L, K, W> C;
  u32* A;
  u32* B;
  GMod<L, K> md;
  u32 lid;

  // A <- A * b * R^-1
  __device__ __forceinline__ void mul(const u32 (&b)[K]) {
    u64 acc[K];
    mont_mul<L, K>(acc, A, b, md);
    u32 x[K];
    seminorm<L, K>(acc, x);
    put_lds<L, K>(A, x, lid);
  }
  // out <- slot * b * R^-1 (slot unchanged)
  __device__ __forceinline__ void mul_to(u32 (&out)[K], const u32* slot, const u32 (&b)[K]) {
    u64 acc[K];
    mont_mul<L, K>(acc, slot, b, md);
    seminorm<L, K>(acc, out);
  }
  __device__ __forceinline__ void mul_global(const u32* __restrict__ src) {
    u32 b[K];
    load_r29<L, K>(b, src, lid);
    mul(b);
  }
  __device__ __forceinline__ void square() {
    u32 b[K];
    get_lds<L, K>(b, A, lid);
    u64 acc[K];
    mont_sqr<L, K>(acc, A, b, md);
    u32 x[K];
    seminorm<L, K>(acc, x);
    put_lds<L, K>(A, x, lid);
  }
  __device__ __forceinline__ void mul_B() {
    u32 b[K];
    get_lds<L, K>(b, B, lid);
    mul(b);
  }
  __device__ __forceinline__ void set(u32* slot, const u32 (&x)[K]) { put_lds<L, K>(slot, x, lid); }
  __device__ __forceinline__ void set_global(u32* slot, const u32* __restrict__ src) {
    u32 x[K];
    load_r29<L, K>(x, src, lid);
    put_lds<L, K>(slot, x, lid);
  }
  __device__ __forceinline__ void save(u32* __restrict__ dst, const u32* slot) {
    u32 x[K];
    get_lds<L, K>(x, slot, lid);
    store_r29<L, K>(dst, x, lid);
  }
  // A W-word number from global memory into this lane's limbs, staged in B
  // (or in A..B when one slot is too small).
  __device__ __forceinline__ void load_words(u32 (&x)[K], const u32* __restrict__ src, int words, int first = 0) {
    u32* st = words <= C::S ? B : A;
    stage_words<L>(st, src, words);
    words_to_limbs<L, K, true>(st, words, first, x);
  }
  // Canonical residue of A written as `words` 32-bit words.
  __device__ __forceinline__ void store_canonical(u32* __restrict__ dst, int words) {
    u32 x[K];
    get_lds<L, K>(x, A, lid);
    canonicalize<L, K>(x, md);
    store_words<L, K>(dst, B, x, words);
  }
  // A <- A * R^-1 (Montgomery form -> Z-space)
  __device__ __forceinline__ void from_mont() {
    u32 b[K];
    one_limbs<L, K>(b, lid);
    mul(b);
  }
};

// This group's view of the block's LDS (modulus staged at lds[0, NL)).
template <int L, int K, int W>
__device__ __forceinline__ Grp<L, K, W> grp_at(u32* lds, const ModArgs& mod) {
  typedef Cfg<L, K, W> C;
  Grp<L, K, W> g;
  g.lid = grp_lane<L>();
  const int gi = (int)(threadIdx.x / (u32)L);
  g.A = lds + C::NL + gi * 2 * C::S;
  g.B = g.A + C::S;
  g.md.m = lds;
  g.md.mprime = mod.mprime;
  return g;
}

// Block prologue: stage the modulus into LDS, then set up this group.
template <int L, int K, int W>
__device__ __forceinline__ Grp<L, K, W> grp_setup(u32* lds, const ModArgs& mod, long& e) {
  typedef Cfg<L, K, W> C;
  for (int i = threadIdx.x; i < C::NL; i += kBlock) lds[i] = mod.m[i];
  __syncthreads();
  e = (long)blockIdx.x * C::kGroups + (int)(threadIdx.x / (u32)L);
  return grp_at<L, K, W>(lds, mod);
}

#define LDS_DECL(L, K, W) __shared__ u32 lds[Cfg<L, K, W>::kLdsWords]

// ---------------------------------------------------------------------------
// Fixed-base table of h_s (public_key.cc:36-39 MakeBaseTable):
// T[w][d] = h_s^(d * 2^(unit*w)) * R mod n^2, R29 interleaved, NL words each.
// Level 0 writes d = 0 (R mod n^2) and d = 1; level j >= 1 writes d in
// (h, 2h], h = 2^(j-1), as T[w][d-h] * T[w][h].
// ---------------------------------------------------------------------------
template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_table_level(ModArgs mod, const u32* __restrict__ bases, u32* table, int windows, int unit,
                                    int level) {
  constexpr int NL = L * K;
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  const long per = 1l << unit;
  if (level == 0) {
    if (e >= windows) return;
    u32 x[K];
    load_r29<L, K>(x, mod.one, g.lid);
    store_r29<L, K>(table + (e * per + 0) * NL, x, g.lid);
    g.load_words(x, bases + e * W, W);
    g.set(g.A, x);
    g.mul_global(mod.r2);
    g.save(table + (e * per + 1) * NL, g.A);
    return;
  }
  const long h = 1l << (level - 1);
  const long w = e / h, d = h + 1 + e % h;
  if (w >= windows || d >= per) return;
  g.set_global(g.A, table + (w * per + (d - h)) * NL);
  g.mul_global(table + (w * per + h) * NL);
  g.save(table + (w * per + d) * NL, g.A);
}

// gm = 1 + n*m mod n^2 for a signed 64-bit m, canonical limbs (encryptor.cc:42-46).
template <int L, int K>
__device__ __forceinline__ void gm_from_i64(u32 (&gm)[K], i64 m, const u32* __restrict__ n29, const GMod<L, K>& md,
                                            u32 lid) {
  const bool neg = m < 0;
  const u64 mag = neg ? (u64)0 - (u64)m : (u64)m;
  const u32 m0 = (u32)mag & kLimbMask, m1 = (u32)(mag >> kLimbBits) & kLimbMask,
            m2 = (u32)(mag >> (2 * kLimbBits));
  i64 v[K];
#pragma unroll
  for (int t = 0; t < K; ++t) {
    const int j = (int)lid * K + t;
    u64 p = (u64)n29[j] * m0;
    if (j >= 1) p += (u64)n29[j - 1] * m1;
    if (j >= 2) p += (u64)n29[j - 2] * m2;
    i64 x = neg ? (i64)md.m[t * L + lid] - (i64)p : (i64)p;
    v[t] = x + (j == 0 ? 1 : 0);
  }
  norm_full<L, K>(v, gm);
}

__device__ __forceinline__ u32 r_digit(const u32* __restrict__ rp, int r_words, int bit, int unit) {
  const int w0 = bit >> 5, s = bit & 31;
  u64 lo = w0 < r_words ? rp[w0] : 0u;
  u64 hi = w0 + 1 < r_words ? rp[w0 + 1] : 0u;
  return (u32)(((hi << 32) | lo) >> s) & ((1u << unit) - 1u);
}

// r must fit the table: bits at and above windows*unit are zero.
__device__ __forceinline__ bool r_fits(const u32* __restrict__ rp, int r_words, int cap_bits) {
  bool ok = true;
  for (int w = 0; w < r_words; ++w) {
    int lo = 32 * w;
    u32 v = rp[w];
    if (lo >= cap_bits) ok &= (v == 0u);
    else if (cap_bits - lo < 32) ok &= (v >> (cap_bits - lo)) == 0u;
  }
  return ok;
}

// A <- h_s^r * R mod n^2 from the table (encryptor.cc:26-31 GetRn).
template <int L, int K, int W>
__device__ __forceinline__ void fixed_base_pow(Grp<L, K, W>& g, const PubArgs& P, const u32* __restrict__ rp,
                                               int r_words) {
  constexpr int NL = L * K;
  const long per = 1l << P.unit;
  g.set_global(g.A, P.table + (long)r_digit(rp, r_words, 0, P.unit) * NL);
  for (int w = 1; w < P.windows; ++w)
    g.mul_global(P.table + (w * per + r_digit(rp, r_words, w * P.unit, P.unit)) * NL);
}

// ---------------------------------------------------------------------------
// Encryption (encryptor.cc:35-49): c = (1 + n*m) * h_s^r mod n^2, Z-space out.
// ---------------------------------------------------------------------------
template <int L, int K, int W>
__device__ __forceinline__ void enc_i64_body(Grp<L, K, W>& g, const PubArgs& P, const i64* __restrict__ m,
                                             const u32* __restrict__ r, int r_words, long e, u32* __restrict__ out,
                                             u32* err) {
  const u32* rp = r + e * r_words;
  if (!r_fits(rp, r_words, P.windows * P.unit)) {
    if (g.lid == 0) atomicOr(err, kErrRange);
    return;
  }
  fixed_base_pow(g, P, rp, r_words);
  u32 gm[K];
  gm_from_i64<L, K>(gm, m[e], P.n29, g.md, g.lid);
  g.mul(gm);  // (h_s^r R) * gm * R^-1 = h_s^r * gm
  g.store_canonical(out + e * W, W);
}

template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_encrypt_i64(PubArgs P, const i64* __restrict__ m,
                                                                              const u32* __restrict__ r, int r_words,
                                                                              long count, u32* __restrict__ out,
                                                                              u32* err) {
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, P.n2, e);
  if (e < count) enc_i64_body(g, P, m, r, r_words, e, out, err);
}

// Encryption of plaintexts given as two's complement over mw words (|m| <= n/2
// is checked by the host).  gm = 1 + n*m mod n^2 via two Montgomery products:
// |m| R (from R^2), then (|m| R) * n * R^-1 = |m| n.
template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_encrypt_big(PubArgs P, const u32* __restrict__ m, int mw, const u32* __restrict__ r,
                                    int r_words, long count, u32* __restrict__ out, u32* err) {
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, P.n2, e);
  if (e >= count) return;
  const u32* rp = r + e * r_words;
  if (!r_fits(rp, r_words, P.windows * P.unit)) {
    if (g.lid == 0) atomicOr(err, kErrRange);
    return;
  }
  // |m| words into B (two's complement negation when the sign bit is set)
  const u32* mp = m + e * mw;
  const bool neg = (mp[mw - 1] >> 31) != 0u;
  wave_lds_sync();
  if (g.lid == 0) {
    u64 c = 1;
    for (int i = 0; i < mw; ++i) {
      u32 v = mp[i];
      if (neg) {
        u64 s = (u64)(u32)~v + c;
        v = (u32)s;
        c = s >> 32;
      }
      g.B[i] = v;
    }
  }
  wave_lds_sync();
  u32 x[K];
  words_to_limbs<L, K>(g.B, mw, 0, x);
  g.set(g.A, x);
  g.mul_global(P.n2.r2);  // |m| R
#pragma unroll
  for (int t = 0; t < K; ++t) x[t] = P.n29[(int)g.lid * K + t];
  g.mul(x);  // |m| n (< 2 n^2)
  get_lds<L, K>(x, g.A, g.lid);
  canonicalize<L, K>(x, g.md);
  i64 v[K];
#pragma unroll
  for (int t = 0; t < K; ++t) {
    const int j = (int)g.lid * K + t;
    i64 y = neg ? (i64)g.md.m[t * L + g.lid] - (i64)x[t] : (i64)x[t];
    v[t] = y + (j == 0 ? 1 : 0);
  }
  norm_full<L, K>(v, x);  // gm = 1 + n m mod n^2 (canonical; m != 0 when negative)
  g.set(g.B, x);
  fixed_base_pow(g, P, rp, r_words);
  g.mul_B();
  g.store_canonical(out + e * W, W);
}

// Randomize (evaluator.cc:25-28) and EncryptZero (encryptor.cc:33):
// out = c * h_s^r mod n^2 (c = 1 when ct is null).
template <int L, int K, int W>
__device__ __forceinline__ void randomize_body(Grp<L, K, W>& g, const PubArgs& P, const u32* __restrict__ ct,
                                               const u32* __restrict__ r, int r_words, long e, u32* __restrict__ out,
                                               u32* err) {
  const u32* rp = r + e * r_words;
  if (!r_fits(rp, r_words, P.windows * P.unit)) {
    if (g.lid == 0) atomicOr(err, kErrRange);
    return;
  }
  u32 c[K];
  if (ct) g.load_words(c, ct + e * W, W);
  else one_limbs<L, K>(c, g.lid);
  g.set(g.B, c);
  fixed_base_pow(g, P, rp, r_words);
  g.mul_B();
  g.store_canonical(out + e * W, W);
}

template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_randomize(PubArgs P, const u32* __restrict__ ct,
                                                                            const u32* __restrict__ r, int r_words,
                                                                            long count, u32* __restrict__ out,
                                                                            u32* err) {
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, P.n2, e);
  if (e < count) randomize_body(g, P, ct, r, r_words, e, out, err);
}

// ---------------------------------------------------------------------------
// ct (+) ct = a*b mod n^2 (evaluator.cc:30-44) and ct (+) pt = c*(1+n*p)
// (evaluator.cc:46-64).  Z-space in and out: (a*b*R^-1) * R^2 * R^-1.
// ---------------------------------------------------------------------------
template <int L, int K, int W>
__device__ __forceinline__ void mulmod_body(Grp<L, K, W>& g, const ModArgs& mod, const u32* __restrict__ a,
                                            const u32* __restrict__ b, long e, u32* __restrict__ out) {
  u32 x[K];
  g.load_words(x, a + e * W, W);
  g.set(g.A, x);
  g.load_words(x, b + e * W, W);
  g.mul(x);
  g.mul_global(mod.r2);
  g.store_canonical(out + e * W, W);
}

template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_mulmod(ModArgs mod, const u32* __restrict__ a,
                                                                         const u32* __restrict__ b, long count,
                                                                         u32* __restrict__ out) {
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e < count) mulmod_body(g, mod, a, b, e, out);
}

template <int L, int K, int W>
__device__ __forceinline__ void add_pt_body(Grp<L, K, W>& g, const PubArgs& P, const u32* __restrict__ ct,
                                            const i64* __restrict__ pt, long e, u32* __restrict__ out) {
  u32 x[K];
  g.load_words(x, ct + e * W, W);
  g.set(g.A, x);
  gm_from_i64<L, K>(x, pt[e], P.n29, g.md, g.lid);
  g.mul(x);
  g.mul_global(P.n2.r2);
  g.store_canonical(out + e * W, W);
}

template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_add_pt_i64(PubArgs P, const u32* __restrict__ ct,
                                                                             const i64* __restrict__ pt, long count,
                                                                             u32* __restrict__ out) {
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, P.n2, e);
  if (e < count) add_pt_body(g, P, ct, pt, e, out);
}

// ---------------------------------------------------------------------------
// Engine form: ciphertexts kept on the device between calls as R29 Montgomery
// limbs (interleaved, NL words per element, c R mod n^2, lazy: below 2 n^2).
// ct (+) ct is then one Montgomery product, (a R)(b R) R^-1 = ab R, read and
// written as limbs with no word conversion; a product of N such values over
// N - 1 products is again in engine form.
// ---------------------------------------------------------------------------
template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_to_mform(ModArgs mod, const u32* __restrict__ ct,
                                                                           long count, u32* __restrict__ out) {
  constexpr int NL = L * K;
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e >= count) return;
  u32 x[K];
  g.load_words(x, ct + e * W, W);
  g.set(g.A, x);
  g.mul_global(mod.r2);  // c R
  g.save(out + e * NL, g.A);
}

template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_mulmod_m(ModArgs mod, const u32* __restrict__ a,
                                                                           const u32* __restrict__ b, long count,
                                                                           u32* __restrict__ out) {
  constexpr int NL = L * K;
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e >= count) return;
  g.set_global(g.A, a + e * NL);
  g.mul_global(b + e * NL);
  g.save(out + e * NL, g.A);
}

// ---------------------------------------------------------------------------
// Powers with 64-bit exponents, fixed 3-bit windows in Montgomery space.
// T (global scratch, room for 2^kPowWin entries of NL words) receives c^d R
// for d < 8.  A holds c R on entry and c^mag R on return.  Leading zero
// windows are skipped per group.  Three bits beat four for the common 32-bit
// magnitudes (6 + 10 products and 30 squarings against 14 + 7 and 28, 8% less
// work at a product's cost of 1.26 squarings) and tie at 64 bits.
// ---------------------------------------------------------------------------
constexpr int kPowWinI64 = 3;
static_assert(kPowWinI64 <= kPowWin, "table room");

template <int L, int K, int W>
__device__ __forceinline__ void pow_u64(Grp<L, K, W>& g, const ModArgs& mod, u64 mag, u32* __restrict__ T) {
  constexpr int NL = L * K;
  constexpr int kW = kPowWinI64;
  constexpr u32 kMask = (1u << kW) - 1u;
  u32 x[K];
  load_r29<L, K>(x, mod.one, g.lid);
  store_r29<L, K>(T, x, g.lid);
  get_lds<L, K>(x, g.A, g.lid);
  store_r29<L, K>(T + NL, x, g.lid);
  g.set(g.B, x);
  for (int d = 2; d < (1 << kW); ++d) {
    g.mul_B();  // A = c^d R
    g.save(T + d * NL, g.A);
  }
  int top = (64 + kW - 1) / kW - 1;  // the top window holds bit 63
  while (top > 0 && ((mag >> (kW * top)) == 0)) --top;
  g.set_global(g.A, T + ((u32)(mag >> (kW * top)) & kMask) * NL);
  for (int w = top - 1; w >= 0; --w) {
#pragma unroll 1
    for (int s = 0; s < kW; ++s) g.square();
    g.mul_global(T + ((u32)(mag >> (kW * w)) & kMask) * NL);
  }
}

// ct (x) scalar = c^k mod n^2 (evaluator.cc:99-123).  k = 0 gives 1, k = 1
// returns c unchanged; for k < 0 this writes c^|k| and sets neg[e] so the
// engine inverts it afterwards.
template <int L, int K, int W>
__device__ __forceinline__ void pow_i64_body(Grp<L, K, W>& g, const ModArgs& mod, const u32* __restrict__ ct,
                                             const i64* __restrict__ k, long e, u32* __restrict__ out, u8* neg,
                                             u32* __restrict__ T) {
  const i64 kk = k[e];
  if (neg && g.lid == 0) neg[e] = kk < 0 ? 1 : 0;
  if (kk == 0 || kk == 1) {
    for (int w = (int)g.lid; w < W; w += L) out[e * W + w] = kk == 0 ? (w == 0 ? 1u : 0u) : ct[e * W + w];
    return;
  }
  const u64 mag = kk < 0 ? (u64)0 - (u64)kk : (u64)kk;
  u32 x[K];
  g.load_words(x, ct + e * W, W);
  g.set(g.A, x);
  g.mul_global(mod.r2);  // c R
  pow_u64(g, mod, mag, T);
  g.from_mont();
  g.store_canonical(out + e * W, W);
}

template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_pow_i64(ModArgs mod, const u32* __restrict__ ct,
                                                                          const i64* __restrict__ k, long count,
                                                                          u32* __restrict__ out, u8* neg, u32* tab) {
  constexpr int NL = L * K;
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e < count) pow_i64_body(g, mod, ct, k, e, out, neg, tab + e * (long)(NL << kPowWin));
}

// ct (x) scalar for exponents of any length: |k| as kw little-endian words
// (kmag) and its sign (kneg, may be null).  Same special cases as k_pow_i64.
template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_pow_words(ModArgs mod, const u32* __restrict__ ct,
                                                                            const u32* __restrict__ kmag,
                                                                            const u8* __restrict__ kneg, int kw,
                                                                            long count, u32* __restrict__ out, u8* neg,
                                                                            u32* tab) {
  constexpr int NL = L * K;
  constexpr u32 kMask = (1u << kPowWin) - 1u;
  constexpr int kPerWord = 32 / kPowWin;
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e >= count) return;
  const u32* kp = kmag + e * kw;
  int topw = kw - 1;
  while (topw >= 0 && kp[topw] == 0u) --topw;
  const bool is_neg = kneg && kneg[e] && topw >= 0;
  if (neg && g.lid == 0) neg[e] = is_neg ? 1 : 0;
  if (topw < 0 || (topw == 0 && kp[0] == 1u && !is_neg)) {
    for (int w = (int)g.lid; w < W; w += L) out[e * W + w] = topw < 0 ? (w == 0 ? 1u : 0u) : ct[e * W + w];
    return;
  }
  u32 x[K];
  g.load_words(x, ct + e * W, W);
  g.set(g.A, x);
  g.mul_global(mod.r2);  // c R
  u32* T = tab + e * (long)(NL << kPowWin);
  load_r29<L, K>(x, mod.one, g.lid);
  store_r29<L, K>(T, x, g.lid);
  get_lds<L, K>(x, g.A, g.lid);
  store_r29<L, K>(T + NL, x, g.lid);
  g.set(g.B, x);
  for (int d = 2; d < (1 << kPowWin); ++d) {
    g.mul_B();
    g.save(T + d * NL, g.A);
  }
  auto digit = [&](int i) { return (kp[i / kPerWord] >> (kPowWin * (i % kPerWord))) & kMask; };
  int top = (topw + 1) * kPerWord - 1;
  while (top > 0 && digit(top) == 0u) --top;
  g.set_global(g.A, T + digit(top) * NL);
  for (int i = top - 1; i >= 0; --i) {
#pragma unroll 1
    for (int s = 0; s < kPowWin; ++s) g.square();
    g.mul_global(T + digit(i) * NL);
  }
  g.from_mont();
  g.store_canonical(out + e * W, W);
}

// ---------------------------------------------------------------------------
// CRT decryption, half h of element e (decryptor.cc:32-53):
//   u_h = c^(p_h - 1) mod p_h^2, canonical, written as W/2 words.
// c is brought into Montgomery space mod p^2 as (c_hi * R^3 + c_lo * R^2) R^-1
// with c = c_hi * 2^(29 NL) + c_lo.  The exponent is shared by all elements;
// its sliding-window schedule (window 5, 16 odd powers) comes from the host,
// so control flow is uniform.
// ---------------------------------------------------------------------------
template <int L, int K, int W>
__device__ __forceinline__ void dec_half_body(Grp<L, K, W>& g, const ModArgs& mod, const u32* __restrict__ sched,
                                              int nsched, const u32* __restrict__ c, u32* __restrict__ T,
                                              u32* __restrict__ dst) {
  constexpr int NL = L * K;
  {
    u32 lo[K], hi[K];
    stage_words<L>(g.A, c, W);  // spans A..B
    words_to_limbs<L, K>(g.A, W, 0, lo);
    words_to_limbs<L, K>(g.A, W, NL, hi);
    g.set(g.A, hi);
    g.set(g.B, lo);
  }
  g.mul_global(mod.r3);  // A = c_hi R^2
  {
    u32 b[K], x[K];
    load_r29<L, K>(b, mod.r2, g.lid);
    g.mul_to(x, g.B, b);  // c_lo R
    get_lds<L, K>(b, g.A, g.lid);
    u64 s[K];
#pragma unroll
    for (int t = 0; t < K; ++t) s[t] = (u64)x[t] + b[t];
    seminorm<L, K>(s, x);  // c R mod p^2 (lazy, < 3 p^2)
    g.set(g.A, x);
  }
  g.save(T, g.A);
  {
    u32 b[K], x2[K];
    get_lds<L, K>(b, g.A, g.lid);
    g.mul_to(x2, g.A, b);  // x^2
    g.set(g.B, x2);
  }
  for (int i = 1; i < (1 << (kDecWin<W> - 1)); ++i) {
    g.mul_B();
    g.save(T + i * NL, g.A);
  }
  // schedule: op = (table index + 1) << 16 | squarings before the multiply
  u32 op = sched[0];
  g.set_global(g.A, T + ((op >> 16) - 1) * NL);
  for (int i = 1; i < nsched; ++i) {
    op = sched[i];
    const u32 sq = op & 0xffffu, idx = op >> 16;
    for (u32 s = 0; s < sq; ++s) g.square();
    if (idx) g.mul_global(T + (idx - 1) * NL);
  }
  g.from_mont();
  g.store_canonical(dst, W / 2);
}

template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_dec_half(DecArgs D, const u32* __restrict__ ct,
                                                                           long count, u32* __restrict__ halves,
                                                                           u32* tab) {
  constexpr int NL = L * K;
  constexpr int HW = W / 2;
  LDS_DECL(L, K, W);
  const int h = blockIdx.y;
  const ModArgs& mod = h ? D.q2 : D.p2;
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e >= count) return;
  dec_half_body(g, mod, h ? D.sched_q : D.sched_p, h ? D.nsched_q : D.nsched_p, ct + e * W,
                tab + (h * count + e) * (long)(NL << (kDecWin<W> - 1)), halves + (h * count + e) * HW);
}

// ---------------------------------------------------------------------------
// CRT recombination, one thread per element, 32-bit words (decryptor.cc:38-52):
//   mp = L_p(u_p) * hp mod p,  mq = L_q(u_q) * hq mod q,
//   m = mp + p * ((mq - mp) * p^-1 mod q),  m -= n if m > n/2.
// L_p(u) = (u - 1) / p is exact: (u - 1) * p^-1 mod 2^(32 PW).
// ---------------------------------------------------------------------------
template <int PW>
__device__ __forceinline__ void mont32(u32 (&r)[PW], const u32 (&a)[PW], const u32* __restrict__ b,
                                       const u32* __restrict__ m, u32 mp) {
  u32 t[PW + 2];
#pragma unroll
  for (int i = 0; i < PW + 2; ++i) t[i] = 0;
  for (int i = 0; i < PW; ++i) {
    u64 c = 0;
    const u32 ai = a[i];
#pragma unroll
    for (int j = 0; j < PW; ++j) {
      u64 v = (u64)ai * b[j] + t[j] + c;
      t[j] = (u32)v;
      c = v >> 32;
    }
    u64 v = (u64)t[PW] + c;
    t[PW] = (u32)v;
    t[PW + 1] = (u32)(v >> 32);
    const u32 q = t[0] * mp;
    v = (u64)q * m[0] + t[0];
    c = v >> 32;
#pragma unroll
    for (int j = 1; j < PW; ++j) {
      v = (u64)q * m[j] + t[j] + c;
      t[j - 1] = (u32)v;
      c = v >> 32;
    }
    v = (u64)t[PW] + c;
    t[PW - 1] = (u32)v;
    t[PW] = t[PW + 1] + (u32)(v >> 32);
  }
  // t < 2m: subtract m once if t >= m
  u32 d[PW];
  i64 br = 0;
#pragma unroll
  for (int j = 0; j < PW; ++j) {
    i64 v = (i64)t[j] - m[j] + br;
    d[j] = (u32)v;
    br = v >> 32;
  }
  const bool ge = (i64)t[PW] + br >= 0;
#pragma unroll
  for (int j = 0; j < PW; ++j) r[j] = ge ? d[j] : t[j];
}

// For a ciphertext coprime to p, u = c^(p-1) mod p^2 is 1 mod p and the
// division is exact.  A value sharing the factor p (not a ciphertext) gives
// u = 0; there (u - 1) / p truncates to 0, as the big-integer division of the
// restated decryptor does, so L = 0.
template <int PW>
__device__ __forceinline__ void l_times_h(u32 (&out)[PW], const u32* __restrict__ u, const u32* __restrict__ pinv,
                                          const u32* __restrict__ hR, const u32* __restrict__ p, u32 pp) {
  u32 nz = 0;
#pragma unroll
  for (int j = 0; j < 2 * PW; ++j) nz |= u[j];
  u32 um1[PW];
  i64 br = nz ? -1 : 0;
#pragma unroll
  for (int j = 0; j < PW; ++j) {
    i64 v = (i64)u[j] + br;
    um1[j] = (u32)v;
    br = v >> 32;
  }
  u32 l[PW];
#pragma unroll
  for (int j = 0; j < PW; ++j) l[j] = 0;
  for (int i = 0; i < PW; ++i) {
    u64 c = 0;
    const u32 ai = um1[i];
    for (int j = 0; i + j < PW; ++j) {
      u64 v = (u64)ai * pinv[j] + l[i + j] + c;
      l[i + j] = (u32)v;
      c = v >> 32;
    }
  }
  mont32<PW>(out, l, hR, p, pp);  // L * h mod p
}

template <int PW>
__global__ __launch_bounds__(kBlock) void k_dec_combine(CombArgs C, const u32* __restrict__ halves, long count,
                                                        u32* __restrict__ out, int out_words, u8* overflow) {
  const long e = (long)blockIdx.x * kBlock + threadIdx.x;
  if (e >= count) return;
  constexpr int HW = 2 * PW;
  u32 mp[PW], mq[PW];
  l_times_h<PW>(mp, halves + e * HW, C.pinv, C.hpR, C.p, C.pprime);
  l_times_h<PW>(mq, halves + (count + e) * HW, C.qinv, C.hqR, C.q, C.qprime);
  // d = mq + 2q - mp in [1, 3q): reduce below q
  u32 d[PW];
  i64 c = 0;
#pragma unroll
  for (int j = 0; j < PW; ++j) {
    i64 v = (i64)mq[j] + 2 * (i64)C.q[j] - (i64)mp[j] + c;
    d[j] = (u32)v;
    c = v >> 32;
  }
  i64 dtop = c;
  for (int rep = 0; rep < 2; ++rep) {
    u32 s[PW];
    i64 b = 0;
#pragma unroll
    for (int j = 0; j < PW; ++j) {
      i64 v = (i64)d[j] - (i64)C.q[j] + b;
      s[j] = (u32)v;
      b = v >> 32;
    }
    if (dtop + b >= 0) {
#pragma unroll
      for (int j = 0; j < PW; ++j) d[j] = s[j];
      dtop += b;
    }
  }
  u32 t[PW];
  mont32<PW>(t, d, C.pinvqR, C.q, C.qprime);  // (mq - mp) p^-1 mod q
  u32 m[2 * PW];
#pragma unroll
  for (int j = 0; j < 2 * PW; ++j) m[j] = j < PW ? mp[j] : 0u;
  for (int i = 0; i < PW; ++i) {
    u64 cc = 0;
    const u32 ti = t[i];
    for (int j = 0; j < PW; ++j) {
      u64 v = (u64)ti * C.p[j] + m[i + j] + cc;
      m[i + j] = (u32)v;
      cc = v >> 32;
    }
    for (int j = i + PW; j < 2 * PW && cc; ++j) {
      u64 v = (u64)m[j] + cc;
      m[j] = (u32)v;
      cc = v >> 32;
    }
  }
  // sign fold: m > n/2 -> m - n (two's complement)
  i64 b = 0;
#pragma unroll
  for (int j = 0; j < 2 * PW; ++j) {
    i64 v = (i64)C.n_half[j] - (i64)m[j] + b;
    b = v >> 32;
  }
  const bool fold = b < 0;
  if (fold) {
    i64 bb = 0;
#pragma unroll
    for (int j = 0; j < 2 * PW; ++j) {
      i64 v = (i64)m[j] - (i64)C.n[j] + bb;
      m[j] = (u32)v;
      bb = v >> 32;
    }
  }
  const u32 ext = fold ? 0xffffffffu : 0u;
  bool fits = true;
  for (int j = out_words; j < 2 * PW; ++j) fits &= (m[j] == ext);
  if (out_words <= 2 * PW) fits &= ((m[out_words - 1] >> 31) == (ext & 1u));
  for (int j = 0; j < out_words; ++j) out[e * out_words + j] = j < 2 * PW ? m[j] : ext;
  if (overflow) overflow[e] = fits ? 0 : 1;
}

// ---------------------------------------------------------------------------
// Segment reductions (numpy/evaluator.cc:355-373 Sum, :209-292 MatMul dot).
// A chunk is a contiguous run of one segment; one group reduces one chunk.
//  - sum: Montgomery products.  Z-space inputs enter as they are, so a chunk's
//    N terms through N-1 products carry R^-(N-1); one product by the
//    Montgomery form of R^N (mod.rpow) leaves every partial in Montgomery
//    form, the form engine-form inputs have already.  Upper levels multiply
//    partials, and k_seg_finish brings each segment's total back to Z-space.
//  - dot (level 0): term = (c R)^|w| in Montgomery space, multiplied into a
//    positive or negative accumulator by the sign of w (w = 0 contributes 1).
// Partials are R29 interleaved.
// ---------------------------------------------------------------------------
// Chunk positions j address element idx[j] when an index list is given
// (gathered sums: selections, bucket sums), else element j.
__device__ __forceinline__ long at_pos(const i64* __restrict__ idx, long j) { return idx ? (long)idx[j] : j; }

// Each partial is the Montgomery form of its chunk's product; an empty chunk
// gives R mod m, the Montgomery form of 1.  Z-space chunks hold at most
// kRpowEntries - 1 terms (the host planners' level-0 chunk).
template <int L, int K, int W, bool ZIN>
__device__ __forceinline__ void reduce_sum_body(Grp<L, K, W>& g, const ModArgs& mod, const u32* __restrict__ in,
                                                const Chunk ch, u32* __restrict__ part,
                                                const i64* __restrict__ idx = nullptr) {
  constexpr int NL = L * K;
  u32 x[K];
  if (ch.len == 0) {
    load_r29<L, K>(x, mod.one, g.lid);
    store_r29<L, K>(part, x, g.lid);
    return;
  }
  if constexpr (ZIN) g.load_words(x, in + at_pos(idx, ch.start) * W, W);
  else load_r29<L, K>(x, in + ch.start * NL, g.lid);
  g.set(g.A, x);
  for (long i = ch.start + 1; i < ch.start + ch.len; ++i) {
    if constexpr (ZIN) g.load_words(x, in + at_pos(idx, i) * W, W);
    else load_r29<L, K>(x, in + i * NL, g.lid);
    g.mul(x);
  }
  if constexpr (ZIN) g.mul_global(mod.rpow + ch.len * NL);  // (prod R^-(len-1)) R^(len+1) R^-1
  g.save(part, g.A);
}

template <int L, int K, int W, bool ZIN>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_reduce_sum(ModArgs mod, const u32* __restrict__ in,
                                                                             const Chunk* __restrict__ chunks,
                                                                             long nchunks, u32* __restrict__ part,
                                                                             const i64* __restrict__ idx) {
  constexpr int NL = L * K;
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e < nchunks) reduce_sum_body<L, K, W, ZIN>(g, mod, in, chunks[e], part + e * NL, idx);
}

template <int L, int K, int W>
__device__ __forceinline__ void reduce_dot0_body(Grp<L, K, W>& g, const ModArgs& mod, const u32* __restrict__ in,
                                                 const i64* __restrict__ wts, const Chunk ch, u32* __restrict__ pp,
                                                 u32* __restrict__ pn, u32* __restrict__ T,
                                                 const i64* __restrict__ idx = nullptr) {
  {
    u32 one[K];
    load_r29<L, K>(one, mod.one, g.lid);
    store_r29<L, K>(pp, one, g.lid);
    store_r29<L, K>(pn, one, g.lid);
  }
  for (long i = ch.start; i < ch.start + ch.len; ++i) {
    const i64 w = wts[i];
    if (w == 0) continue;
    const u64 mag = w < 0 ? (u64)0 - (u64)w : (u64)w;
    u32 x[K];
    g.load_words(x, in + at_pos(idx, i) * W, W);
    g.set(g.A, x);
    g.mul_global(mod.r2);  // c R
    if (mag != 1) pow_u64(g, mod, mag, T);
    // accumulator <- accumulator * term; accumulators live in global memory
    u32* acc = w > 0 ? pp : pn;
    get_lds<L, K>(x, g.A, g.lid);
    g.set_global(g.A, acc);
    g.mul(x);
    g.save(acc, g.A);
  }
}

template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_reduce_dot0(ModArgs mod, const u32* __restrict__ in,
                                                                              const i64* __restrict__ wts,
                                                                              const Chunk* __restrict__ chunks,
                                                                              long nchunks, u32* __restrict__ part_pos,
                                                                              u32* __restrict__ part_neg, u32* tab,
                                                                              const i64* __restrict__ idx) {
  constexpr int NL = L * K;
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e < nchunks)
    reduce_dot0_body(g, mod, in, wts, chunks[e], part_pos + e * NL, part_neg + e * NL,
                     tab + e * (long)(NL << kPowWin), idx);
}

// Upper levels of a dot: products of Montgomery-form partials (stay in M-space).
template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_reduce_mont(ModArgs mod, const u32* __restrict__ in, const Chunk* __restrict__ chunks,
                                    long nchunks, u32* __restrict__ part) {
  constexpr int NL = L * K;
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e >= nchunks) return;
  const Chunk ch = chunks[e];
  g.set_global(g.A, mod.one);
  for (long i = ch.start; i < ch.start + ch.len; ++i) g.mul_global(in + i * NL);
  g.save(part + e * NL, g.A);
}

// Finish a segment sum (or a dot's negative product, or engine-form values
// leaving the engine): Montgomery form -> canonical Z-space words.
template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_seg_finish(ModArgs mod, const u32* __restrict__ part,
                                                                             long nseg, u32* __restrict__ out) {
  constexpr int NL = L * K;
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e >= nseg) return;
  g.set_global(g.A, part + e * NL);
  g.from_mont();
  g.store_canonical(out + e * W, W);
}

// out = x * y mod m where x is Montgomery form (R29) and y is Z-space words:
// x R * y * R^-1 = x y.  Finishes a dot as P * N^-1.
template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_mont_times_z(ModArgs mod, const u32* __restrict__ xm, const u32* __restrict__ yz, long count,
                                     u32* __restrict__ out) {
  constexpr int NL = L * K;
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e >= count) return;
  u32 y[K];
  g.load_words(y, yz + e * W, W);
  g.set_global(g.A, xm + e * NL);
  g.mul(y);
  g.store_canonical(out + e * W, W);
}

// ---------------------------------------------------------------------------
// Inclusive prefix products within segments (cumulative bucket sums,
// numpy/evaluator.cc:450-456): out[j] = in[start] * ... * in[j] mod m for j
// in each segment, Z-space in and out; one group walks one segment.  Each step
// is (acc * x) R^-1 * R^2 R^-1.
// ---------------------------------------------------------------------------
template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_prefix_mul(ModArgs mod, const u32* __restrict__ in,
                                                                             const Chunk* __restrict__ segs, long nseg,
                                                                             u32* __restrict__ out) {
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e >= nseg) return;
  const Chunk sg = segs[e];
  if (sg.len <= 0) return;
  u32 x[K];
  g.load_words(x, in + sg.start * W, W);
  g.set(g.A, x);
  g.store_canonical(out + sg.start * W, W);
  for (long j = sg.start + 1; j < sg.start + sg.len; ++j) {
    g.load_words(x, in + j * W, W);
    g.mul(x);
    g.mul_global(mod.r2);
    g.store_canonical(out + j * W, W);
  }
}

// ---------------------------------------------------------------------------
// Batch inversion (Montgomery's trick) for Negate / Sub / negative scalars
// (evaluator.cc:66-97).  Level-0 inputs are Z-space words selected by `sel`
// (unselected elements count as 1 and are copied through); upper levels work
// on Montgomery-form R29 values.
//   prefix: per chunk, P_j = X_0 ... X_j (Montgomery form), chunk total out;
//   down:   given the inverse of the chunk total, y_j = inv * P_(j-1),
//           inv = inv * X_j, y_0 = inv.
// ---------------------------------------------------------------------------
// x <- X_i = x_i R (Montgomery form of element i), via slot B.
template <int L, int K, int W, bool ZIN>
__device__ __forceinline__ bool inv_x(u32 (&x)[K], Grp<L, K, W>& g, const ModArgs& mod, const u32* in,
                                      const u8* __restrict__ sel, long i) {
  constexpr int NL = L * K;
  if constexpr (ZIN) {
    if (sel && !sel[i]) return false;
    g.load_words(x, in + i * W, W);
    g.set(g.B, x);
    u32 b[K];
    load_r29<L, K>(b, mod.r2, g.lid);
    g.mul_to(x, g.B, b);
  } else {
    load_r29<L, K>(x, in + i * NL, g.lid);
  }
  return true;
}

template <bool ZIN>
__device__ __forceinline__ bool chunk_selected(const u8* __restrict__ sel, long first, long last) {
  if (!ZIN || !sel) return true;
  bool any = false;
  for (long i = first; i < last; ++i) any |= sel[i] != 0;
  return any;
}

// Level-0 selections by index list: the selected elements' positions, in any
// order (each element's inverse does not depend on how elements are grouped),
// packed with one atomic per wave.  *nsel must be 0 on entry.  (Templates
// only so that every translation unit including this header may define them.)
template <int kUnused = 0>
__global__ __launch_bounds__(kBlock) void k_inv_compact(const u8* __restrict__ sel, long count, u32* __restrict__ idx,
                                                         u32* nsel) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  const bool take = i < count && sel[i] != 0;
  const u64 mask = __builtin_amdgcn_ballot_w64(take);
  if (mask == 0) return;  // uniform
  const u32 lane = threadIdx.x & 63u;
  const u32 first = (u32)__builtin_ctzll(mask);
  u32 base = 0;
  if (lane == first) base = atomicAdd(nsel, (u32)__builtin_popcountll(mask));
  base = (u32)__shfl((int)base, (int)first, 64);
  if (take) idx[base + (u32)__builtin_popcountll(mask & ((1ull << lane) - 1ull))] = (u32)i;
}

template <int kUnused = 0>
__global__ __launch_bounds__(64) void k_zero_u32(u32* p) {
  if (threadIdx.x == 0) p[0] = 0u;
}

// Position j of level 0 -> element index; last position that holds one.
__device__ __forceinline__ long inv_elem(const InvSel& s, long j) { return s.idx ? (long)s.idx[j] : j; }
__device__ __forceinline__ long inv_last(const InvSel& s, long last) {
  if (!s.idx) return last;
  const long n = (long)*s.nsel;
  return n < last ? n : last;
}

template <int L, int K, int W, bool ZIN>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_inv_prefix(ModArgs mod, const u32* __restrict__ in, InvSel s, long count,
                                   int chunk, u32* __restrict__ prefix, u32* __restrict__ totals) {
  constexpr int NL = L * K;
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  const long first = e * chunk;
  if (first >= count) return;
  long last = first + chunk < count ? first + chunk : count;
  if constexpr (ZIN) last = inv_last(s, last);  // positions past the list count as 1
  g.set_global(g.A, mod.one);
  if (first < last && chunk_selected<ZIN>(s.sel, first, last)) {
    for (long j = first; j < last; ++j) {
      u32 x[K];
      const long i = ZIN ? inv_elem(s, j) : j;
      if (inv_x<L, K, W, ZIN>(x, g, mod, in, s.sel, i)) g.mul(x);  // unselected: X = 1
      g.save(prefix + j * NL, g.A);
    }
  }
  g.save(totals + e * NL, g.A);
}

// Down pass.  `out` may alias `in`: element i is read (to rebuild X_i) before
// out[i] is written.  y at position j is parked in the spent prefix slot
// P_(j-1).  With an index list the inversion is in place (out == in) and
// elements off the list are left as they are.
template <int L, int K, int W, bool ZIN>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_inv_down(ModArgs mod, const u32* in, InvSel s,
                                                                           long count, int chunk, u32* prefix,
                                                                           const u32* __restrict__ tot_inv, u32* out) {
  constexpr int NL = L * K;
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  const long first = e * chunk;
  if (first >= count) return;
  long last = first + chunk < count ? first + chunk : count;
  if constexpr (ZIN) {
    if (s.idx) {
      last = inv_last(s, last);
      if (first >= last) return;
    }
  }
  const u8* sel = s.sel;
  if (!chunk_selected<ZIN>(sel, first, last)) {
    if constexpr (ZIN) {
      if (out != in)
        for (long i = first; i < last; ++i)
          for (int w = (int)g.lid; w < W; w += L) out[i * W + w] = in[i * W + w];
    }
    return;
  }
  g.set_global(g.A, tot_inv + e * NL);  // A = inv (Montgomery form)
  for (long j = last - 1; j >= first; --j) {
    const long i = ZIN ? inv_elem(s, j) : j;
    if constexpr (ZIN) {
      if (sel && !sel[i]) {
        if (out != in)
          for (int w = (int)g.lid; w < W; w += L) out[i * W + w] = in[i * W + w];
        continue;
      }
    }
    u32* ypark = nullptr;
    if (j > first) {
      u32 x[K];
      inv_x<L, K, W, ZIN>(x, g, mod, in, sel, i);  // X_i, read before out[i] is written
      g.set(g.B, x);
      ypark = prefix + (j - 1) * NL;
      u32 b[K];
      load_r29<L, K>(b, ypark, g.lid);
      g.mul_to(x, g.A, b);  // y_i = X_i^-1 R = inv * P_(i-1)
      store_r29<L, K>(ypark, x, g.lid);
      g.mul_B();  // inv <- inv * X_i
    }
    u32 y[K];
    if (ypark) load_r29<L, K>(y, ypark, g.lid);
    else get_lds<L, K>(y, g.A, g.lid);
    if constexpr (ZIN) {
      // to Z-space and out; B is the operand and then the staging area
      g.set(g.B, y);
      u32 one[K];
      one_limbs<L, K>(one, g.lid);
      g.mul_to(y, g.B, one);
      canonicalize<L, K>(y, g.md);
      store_words<L, K>(out + i * W, g.B, y, W);
    } else {
      store_r29<L, K>(out + i * NL, y, g.lid);
    }
  }
}

// ---------------------------------------------------------------------------
// The one inversion per batch: x^-1 mod m by Bernstein-Yang divsteps (D. J.
// Bernstein, B.-Y. Yang, "Fast constant-time gcd computation and modular
// inversion", 2019), 60 divsteps at a time.  Wave 1 holds f, g (the gcd pair,
// starting at m and x) and d, e (f = d x, g = e x mod m) as 128 signed limbs of
// 60 bits, two per lane.  A batch runs 60 divsteps on the low 64 bits of f and
// g (uniform values), which yield the transition matrix [u v; q r] with
// |u| + |v|, |q| + |r| <= 2^60; then f, g <- [u v; q r] (f, g) / 2^60 and d, e
// likewise after adding the multiple k m of m that makes them divisible.  Wave
// 0 runs the divsteps one batch ahead: from the exact low 180 bits of f and g
// before a batch and that batch's matrix it has the low bits after it, so the
// next matrix is ready when wave 1 finishes applying the current one.  Limbs stay redundant, in [-2, 2^60 + 1] with the top one whole,
// so they are the exact integers and no carry crosses more than one lane per
// round; one lane resolves them to check for the end.  After the published
// bound of (49 N + 57) / 17 divsteps (N = bits of m) g = 0 and f = +-gcd(x, m),
// and typical inputs get there after about 72% of it; the kernel checks from
// 70% on and runs further batches past the bound if it ever had to.
// ---------------------------------------------------------------------------
constexpr int kSgBits = 60;
constexpr u64 kSgMask = (1ull << kSgBits) - 1ull;
constexpr int kSgLimbs = 128;  // 7680 bits: |values| stay below 2^(N + 10)

struct SgNum {
  i64 l[2];  // limbs 2 lane, 2 lane + 1
};

__device__ __forceinline__ i64 sg_first64(i64 x) {  // lane 0's value, in every lane
  const u32 lo = __builtin_amdgcn_readfirstlane((u32)(u64)x), hi = __builtin_amdgcn_readfirstlane((u32)((u64)x >> 32));
  return (i64)(((u64)hi << 32) | lo);
}

// The low 64 bits of the exact integer (limb 0 and the low bits of limb 1).
__device__ __forceinline__ u64 sg_low64(const SgNum& x) {
  return (u64)sg_first64(x.l[0]) + ((u64)sg_first64(x.l[1]) << kSgBits);
}

// The exact low 128 bits of a number from its three lowest limbs, in every lane.
__device__ __forceinline__ unsigned __int128 sg_low128(const i64* l3) {
  typedef unsigned __int128 u128;
  return (u128)(__int128)sg_first64(l3[0]) + ((u128)(__int128)sg_first64(l3[1]) << kSgBits) +
         ((u128)(__int128)sg_first64(l3[2]) << (2 * kSgBits));
}

// 60 divsteps (delta, f, g) -> (1 - delta, g, (g - f) / 2) if delta > 0 and g
// is odd, else (1 + delta, f, (g + (g mod 2) f) / 2), on the low 64 bits, with
// the matrix such that f' 2^60 = u f + v g and g' 2^60 = q f + r g.  Several
// divsteps are taken per iteration (the variable-time form of the method, as
// in P. Wuille's safegcd notes for libsecp256k1): a run of z even steps at
// once (z = trailing zeros of g: g >>= z, u, v <<= z), then, with g odd, the
// next `limit` steps, which cannot swap while delta stays <= 0, as one
// elimination g += w f with w = -g / f mod 2^limit (limit <= 6; 1 / f mod 64
// is f (2 - f^2) for odd f).  A swap step first exchanges the rows as
// (f, g) <- (g, -f).  The result is the same matrix and delta as one divstep
// at a time, in about half the iterations.
__device__ __forceinline__ int sg_divsteps(int delta, u64 f, u64 g, i64& u, i64& v, i64& q, i64& r) {
  u64 uu = 1, vv = 0, qq = 0, rr = 1;
  int i = kSgBits;
  for (;;) {
    const int z = __builtin_ctzll(g | (~0ull << i));  // at most i
    g >>= z;
    uu <<= z;
    vv <<= z;
    delta += z;
    i -= z;
    if (i == 0) break;
    if (delta > 0) {  // g odd: swap
      delta = -delta;
      u64 t = f;
      f = g;
      g = 0u - t;
      t = uu;
      uu = qq;
      qq = 0u - t;
      t = vv;
      vv = rr;
      rr = 0u - t;
    }
    const int limit = 1 - delta < i ? 1 - delta : i;  // >= 1
    const u32 m = (u32)((~0ull >> (64 - limit)) & 63u);
    const u32 f32 = (u32)f, w = (f32 * (u32)g * (f32 * f32 - 2u)) & m;
    g += f * w;
    qq += uu * w;
    rr += vv * w;
  }
  u = (i64)uu;
  v = (i64)vv;
  q = (i64)qq;
  r = (i64)rr;
  return delta;
}

// out = (a x + b y + k m) / 2^60 (exact); out may be x or y.
__device__ __forceinline__ void sg_lincomb(SgNum& out, i64 a, i64 b, const SgNum& x, const SgNum& y, u64 k,
                                           const u64 (&ml)[2], u32 lane) {
  __int128 P[2];
#pragma unroll
  for (int j = 0; j < 2; ++j)
    P[j] = (__int128)a * x.l[j] + (__int128)b * y.l[j] + (__int128)((unsigned __int128)k * ml[j]);
  const bool top = lane == 63u;
  const i64 lo0 = (i64)((u64)P[0] & kSgMask), hi0 = (i64)(P[0] >> kSgBits);
  const i64 lo1 = (i64)((u64)P[1] & kSgMask), hi1 = (i64)(P[1] >> kSgBits);
  i64 below = __shfl_up(hi1, 1, 64);
  if (lane == 0u) below = 0;
  const i64 t0 = lo0 + below;
  const i64 t1 = top ? (i64)P[1] + hi0 : lo1 + hi0;  // the top limb is never split
  const i64 hb0 = t0 >> kSgBits, hb1 = t1 >> kSgBits;
  i64 below2 = __shfl_up(hb1, 1, 64);
  if (lane == 0u) below2 = 0;
  const i64 s0 = (t0 & (i64)kSgMask) + below2;
  const i64 s1 = top ? t1 + hb0 : (t1 & (i64)kSgMask) + hb0;
  i64 up = __shfl_down(s0, 1, 64);  // limb 0 of the result is 0: shift down one limb
  if (top) up = 0;
  out.l[0] = s1;
  out.l[1] = up;
}

// Limb `li` (60 bits) of a little-endian word array; words past nw are 0.
__device__ __forceinline__ u64 sg_limb_of_words(const u32* w, int nw, int li) {
  const int bit = li * kSgBits, w0 = bit >> 5, sh = bit & 31;
  const u64 a = w0 < nw ? w[w0] : 0u, b = w0 + 1 < nw ? w[w0 + 1] : 0u, c = w0 + 2 < nw ? w[w0 + 2] : 0u;
  u64 v = (a | (b << 32)) >> sh;
  if (sh) v |= c << (64 - sh);
  return v & kSgMask;
}

// One lane, in place: redundant limbs -> limbs in [0, 2^60); returns the carry
// out (0 or -1: the sign of the exact integer).
__device__ __forceinline__ i64 sg_resolve(i64* io) {
  i64 c = 0;
  for (int i = 0; i < kSgLimbs; ++i) {
    const i64 v = io[i] + c;
    io[i] = v & (i64)kSgMask;
    c = v >> kSgBits;
  }
  return c;
}

// Inverse of one Montgomery-form value (R29), in place: a group of L lanes maps
// it to canonical Z-space words, the whole wave runs the divsteps, one lane
// resolves the result, and the group maps it back (x R -> x^-1 R).  Sets
// kErrNotInvertible if gcd(x, m) != 1.  One wave, launched once per batch.
template <int L, int K, int W>
__global__ __launch_bounds__(128) void k_inv_single(ModArgs mod, const u32* __restrict__ m_words, u32* xio, u32* err) {
  constexpr int NL = L * K;
  static_assert(kSgBits * kSgLimbs >= 32 * W + 64, "room for the divsteps numbers");
  static_assert(29 * NL >= 32 * (W + 1), "room for the result before reduction");
  __shared__ u32 lds[NL + 2 * Cfg<L, K, W>::S];
  __shared__ u32 words[W + 2];
  __shared__ i64 res[3][kSgLimbs];
  __shared__ int s_state;  // 0 running, 1 done (+2: negate), -1 not invertible
  for (int i = threadIdx.x; i < NL; i += 128) lds[i] = mod.m[i];
  __syncthreads();
  Grp<L, K, W> g;
  g.lid = grp_lane<L>();
  g.A = lds + NL;
  g.B = g.A + Cfg<L, K, W>::S;
  g.md.m = lds;
  g.md.mprime = mod.mprime;
  const bool in_grp = threadIdx.x < (u32)L;
  if (in_grp) {
    g.set_global(g.A, xio);
    g.from_mont();
    u32 x[K];
    get_lds<L, K>(x, g.A, g.lid);
    canonicalize<L, K>(x, g.md);
    g.set(g.A, x);
    for (int w = (int)g.lid; w < W; w += L) words[w] = limbs_to_word<L, K>(g.A, w);
  }
  __syncthreads();
  // Wave 0 runs the divsteps, one batch ahead; wave 1 holds the numbers and
  // applies each batch's matrix.  Per batch, wave 1 publishes the three lowest
  // limbs of f and g (their exact low 180 bits), and wave 0 publishes the next
  // batch's matrix, in buffers alternating by batch parity.
  const u32 lane = threadIdx.x & 63u;
  const bool steps_wave = threadIdx.x < 64u;
  __shared__ i64 s_low[2][2][3];  // [parity][f, g][limb 0..2]
  __shared__ i64 s_mat[2][4];     // [parity][u, v, q, r]
  SgNum f, gg, d, e;
  u64 ml[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    ml[j] = sg_limb_of_words(m_words, W, (int)(2 * lane + j));
    f.l[j] = (i64)ml[j];
    gg.l[j] = (i64)sg_limb_of_words(words, W, (int)(2 * lane + j));
    d.l[j] = 0;
    e.l[j] = (lane == 0u && j == 0) ? 1 : 0;
  }
  auto publish_low = [&](int buf) {
    if (lane == 0u) {
      s_low[buf][0][0] = f.l[0];
      s_low[buf][0][1] = f.l[1];
      s_low[buf][1][0] = gg.l[0];
      s_low[buf][1][1] = gg.l[1];
    } else if (lane == 1u) {
      s_low[buf][0][2] = f.l[0];
      s_low[buf][1][2] = gg.l[0];
    }
  };
  // m^-1 mod 2^64 by Newton's iteration (m odd: m m = 1 mod 8 to start)
  const u64 m0 = (u64)m_words[0] | ((u64)m_words[1] << 32);
  u64 minv = m0;
#pragma unroll
  for (int i = 0; i < 5; ++i) minv *= 2u - m0 * minv;
  const int nbits = 32 * W;
  const int batches = ((49 * nbits + 57) / 17 + kSgBits) / kSgBits;
  // Random inputs reach g = 0 after 71-73% of the bound (4096 and 6144 bits),
  // and once g = 0 further batches leave f and d unchanged: check from 70% of
  // the bound, every 4 batches up to it, every 8 after it.
  const int first = batches * 7 / 10;
  int delta = 1, done = 0, ph = 0;
  i64 mu = 0, mv = 0, mq = 0, mr = 0;  // wave 0: the matrix of the batch wave 1 applies next
  if (!steps_wave) publish_low(0);
  __syncthreads();
  if (steps_wave) {
    delta = sg_divsteps(delta, (u64)sg_low128(s_low[0][0]), (u64)sg_low128(s_low[0][1]), mu, mv, mq, mr);
    if (lane == 0u) {
      s_mat[0][0] = mu;
      s_mat[0][1] = mv;
      s_mat[0][2] = mq;
      s_mat[0][3] = mr;
    }
  }
  __syncthreads();
  for (;;) {
    const int todo = done == 0 ? first : done < batches ? 4 : 8;
    for (int bi = 0; bi < todo; ++bi, ++ph) {
      const int cur = ph & 1, nxt = cur ^ 1;
      if (steps_wave) {
        // low bits of the numbers after this batch, then the next batch's matrix
        typedef unsigned __int128 u128;
        const u128 F = sg_low128(s_low[cur][0]), G = sg_low128(s_low[cur][1]);
        const u128 F1 = (u128)(__int128)mu * F + (u128)(__int128)mv * G;
        const u128 G1 = (u128)(__int128)mq * F + (u128)(__int128)mr * G;
        delta = sg_divsteps(delta, (u64)(F1 >> kSgBits), (u64)(G1 >> kSgBits), mu, mv, mq, mr);
        if (lane == 0u) {
          s_mat[nxt][0] = mu;
          s_mat[nxt][1] = mv;
          s_mat[nxt][2] = mq;
          s_mat[nxt][3] = mr;
        }
      } else {
        const i64 u = sg_first64(s_mat[cur][0]), v = sg_first64(s_mat[cur][1]);
        const i64 q = sg_first64(s_mat[cur][2]), r = sg_first64(s_mat[cur][3]);
        const u64 d0 = (u64)sg_first64(d.l[0]), e0 = (u64)sg_first64(e.l[0]);
        const u64 kd = (0u - ((u64)u * d0 + (u64)v * e0) * minv) & kSgMask;
        const u64 ke = (0u - ((u64)q * d0 + (u64)r * e0) * minv) & kSgMask;
        const u64 none[2] = {0u, 0u};
        SgNum nf, ng, nd, ne;
        sg_lincomb(nf, u, v, f, gg, 0u, none, lane);
        sg_lincomb(ng, q, r, f, gg, 0u, none, lane);
        sg_lincomb(nd, u, v, d, e, kd, ml, lane);
        sg_lincomb(ne, q, r, d, e, ke, ml, lane);
        f = nf;
        gg = ng;
        d = nd;
        e = ne;
        publish_low(nxt);
      }
      __syncthreads();
    }
    done += todo;
    // resolve g (then f and d) in thread 0
    if (!steps_wave) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        res[0][2 * lane + j] = f.l[j];
        res[1][2 * lane + j] = gg.l[j];
        res[2][2 * lane + j] = d.l[j];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0u) {
      i64 *rf = res[0], *rg = res[1], *rd = res[2];
      bool g_zero = sg_resolve(rg) == 0;
      for (int i = 0; i < kSgLimbs && g_zero; ++i) g_zero = rg[i] == 0;
      bool f_one = false, f_mone = false;
      if (g_zero) {
        const i64 cf = sg_resolve(rf);
        f_one = cf == 0 && rf[0] == 1;
        f_mone = cf == -1 && rf[0] == (i64)kSgMask;
        for (int i = 1; i < kSgLimbs; ++i) {
          f_one &= rf[i] == 0;
          f_mone &= rf[i] == (i64)kSgMask;
        }
      }
      if (g_zero && (f_one || f_mone)) {
        // |d| as W + 1 words; the sign of d f into the state
        const bool dneg = sg_resolve(rd) < 0;
        if (dneg) {
          i64 c = 1;
          for (int i = 0; i < kSgLimbs; ++i) {
            const i64 t = ((~rd[i]) & (i64)kSgMask) + c;
            rd[i] = t & (i64)kSgMask;
            c = t >> kSgBits;
          }
        }
        for (int w = 0; w < W + 1; ++w) {
          const int bit = 32 * w, li = bit / kSgBits, sh = bit - kSgBits * li;
          u64 v = (u64)rd[li] >> sh;
          if (li + 1 < kSgLimbs && sh > kSgBits - 32) v |= (u64)rd[li + 1] << (kSgBits - sh);
          words[w] = (u32)v;
        }
        s_state = 1 + ((dneg != f_mone) ? 2 : 0);
      } else if (g_zero) {
        s_state = -1;  // f = +-gcd(x, m) != +-1
      } else {
        s_state = done >= 4 * batches ? -1 : 0;
      }
    }
    __syncthreads();
    if (s_state != 0) break;  // uniform
  }
  if (s_state < 0) {
    if (threadIdx.x == 0u) atomicOr(err, kErrNotInvertible);
    return;  // uniform
  }
  const bool neg = (s_state & 2) != 0;
  __syncthreads();
  if (in_grp) {
    u32 x[K];
    words_to_limbs<L, K>(words, W + 1, 0, x);
    g.set(g.A, x);
    g.mul_global(mod.r2);  // |d| R mod m (|d| < R)
    if (neg) {
      get_lds<L, K>(x, g.A, g.lid);
      canonicalize<L, K>(x, g.md);
      i64 t[K];
#pragma unroll
      for (int i = 0; i < K; ++i) t[i] = (i64)g.md.m[i * L + g.lid] - (i64)x[i];
      u32 y[K];
      norm_full<L, K>(t, y);  // m - |d| R: in [1, m) since the value is invertible
      g.set(g.A, y);
    }
    g.save(xio, g.A);
  }
}

// Montgomery-form export (encryptor.cc:46 MapIntoMSpace) for R_x = 2^rbits:
// out = c * cz mod m with cz = R_x mod m (or R_x^-1 mod m for the import),
// computed as mont(mont(c, cz), R^2).
template <int L, int K, int W>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_mul_const(ModArgs mod, const u32* __restrict__ in, const u32* __restrict__ cz, long count,
                                  u32* __restrict__ out) {
  LDS_DECL(L, K, W);
  long e;
  auto g = grp_setup<L, K, W>(lds, mod, e);
  if (e >= count) return;
  u32 x[K];
  g.load_words(x, in + e * W, W);
  g.set(g.A, x);
  g.load_words(x, cz, W);
  g.mul(x);
  g.mul_global(mod.r2);
  g.store_canonical(out + e * W, W);
}

// ---------------------------------------------------------------------------
// Persistent work-list kernel.  Workgroups pull units (blocks of kGroups
// elements of one item) from ctr[0] until the list is exhausted, so one launch
// covers vectors of any length and any key of this shape, and long and short
// items balance across the chip.  Window tables are per group slot
// (blockIdx, group), reused from unit to unit.  kDec selects the decryption
// half kernel (moduli p^2 / q^2); otherwise the n^2 ops.  The last workgroup to
// finish clears the counters, so launches need no memset between them (and a
// captured graph can replay the launch as is).
// ---------------------------------------------------------------------------
template <int L, int K, int W, bool kDec>
__global__ __launch_bounds__(kBlock, (Cfg<L, K, W>::kWaves)) void k_worklist(const WorkItem* __restrict__ items,
                                                                           int nitems, u32 total_units, u32* ctr,
                                                                           u32* slot_tab) {
  typedef Cfg<L, K, W> C;
  constexpr int NL = L * K;
  LDS_DECL(L, K, W);
  __shared__ u32 s_unit;
  const int gi = (int)(threadIdx.x / (u32)L);
  const long slot = (long)blockIdx.x * C::kGroups + gi;
  for (;;) {
    if (threadIdx.x == 0) s_unit = atomicAdd(ctr, 1u);
    __syncthreads();
    // uniform by construction; readfirstlane lets the item lookup and every
    // item field live in scalar registers
    const u32 u = __builtin_amdgcn_readfirstlane(s_unit);
    if (u >= total_units) break;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (items[mid].unit0 <= (long)u) lo = mid;
      else hi = mid - 1;
    }
    const WorkItem& it = items[lo];
    for (int i = threadIdx.x; i < NL; i += kBlock) lds[i] = it.pub.n2.m[i];
    __syncthreads();
    auto g = grp_at<L, K, W>(lds, it.pub.n2);
    const long e = ((long)u - it.unit0) * C::kGroups + gi;
    if (e < it.count) {
      if constexpr (kDec) {
        dec_half_body(g, it.pub.n2, static_cast<const u32*>(it.c), it.aux, static_cast<const u32*>(it.a) + e * W,
                      slot_tab + slot * (long)(NL << (kDecWin<W> - 1)), static_cast<u32*>(it.out) + e * (W / 2));
      } else {
        u32* T = slot_tab + slot * (long)(NL << kPowWin);
        switch (it.op) {
          case kWorkEncrypt:
            enc_i64_body(g, it.pub, static_cast<const i64*>(it.a), static_cast<const u32*>(it.b), it.aux, e,
                         static_cast<u32*>(it.out), it.err);
            break;
          case kWorkRandomize:
            randomize_body(g, it.pub, static_cast<const u32*>(it.a), static_cast<const u32*>(it.b), it.aux, e,
                           static_cast<u32*>(it.out), it.err);
            break;
          case kWorkAdd:
            mulmod_body(g, it.pub.n2, static_cast<const u32*>(it.a), static_cast<const u32*>(it.b), e,
                        static_cast<u32*>(it.out));
            break;
          case kWorkAddPt:
            add_pt_body(g, it.pub, static_cast<const u32*>(it.a), static_cast<const i64*>(it.b), e,
                        static_cast<u32*>(it.out));
            break;
          case kWorkMul:
            pow_i64_body(g, it.pub.n2, static_cast<const u32*>(it.a), static_cast<const i64*>(it.b), e,
                         static_cast<u32*>(it.out), it.neg, T);
            break;
          case kWorkDot0:
            reduce_dot0_body(g, it.pub.n2, static_cast<const u32*>(it.a), static_cast<const i64*>(it.b),
                             static_cast<const Chunk*>(it.c)[e], static_cast<u32*>(it.out) + e * NL,
                             static_cast<u32*>(it.out2) + e * NL, T, static_cast<const i64*>(it.d));
            break;
          case kWorkSum0:
            reduce_sum_body<L, K, W, true>(g, it.pub.n2, static_cast<const u32*>(it.a),
                                           static_cast<const Chunk*>(it.c)[e], static_cast<u32*>(it.out) + e * NL,
                                           static_cast<const i64*>(it.d));
            break;
          default:
            break;
        }
      }
    }
    __syncthreads();  // LDS (modulus, slots, s_unit) is rewritten by the next unit
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(ctr + 1, 1u) == gridDim.x - 1) {
      atomicExch(ctr, 0u);
      atomicExch(ctr + 1, 0u);
    }
  }
}

// ---------------------------------------------------------------------------
// Host-side launchers, one instantiation per supported shape
// ---------------------------------------------------------------------------
template <int L, int K, int W>
struct Shape {
  static constexpr int kL = L, kK = K, kW = W, kNL = L * K;
};

static inline unsigned blocks_for(long groups, int L) {
  const long per = kBlock / L;
  return (unsigned)((groups + per - 1) / per);
}

#define HEU_LAUNCH(kern, groups, L, stream, ...)                                                   \
  do {                                                                                             \
    if ((groups) > 0) hipLaunchKernelGGL(kern, dim3(blocks_for((groups), (L))), dim3(kBlock), 0, \
                                         (stream), __VA_ARGS__);                                   \
  } while (0)

template <class S>
struct Launch {
  static constexpr int L = S::kL, K = S::kK, W = S::kW;
  static void table_level(hipStream_t st, const ModArgs& mod, const u32* bases, u32* table, int windows, int unit,
                          int level) {
    long groups = level == 0 ? windows : (long)windows * (1l << (level - 1));
    HEU_LAUNCH((k_table_level<L, K, W>), groups, L, st, mod, bases, table, windows, unit, level);
  }
  static void encrypt_i64(hipStream_t st, const PubArgs& P, const i64* m, const u32* r, int rw, long n, u32* out,
                          u32* err) {
    HEU_LAUNCH((k_encrypt_i64<L, K, W>), n, L, st, P, m, r, rw, n, out, err);
  }
  static void encrypt_big(hipStream_t st, const PubArgs& P, const u32* m, int mw, const u32* r, int rw, long n,
                          u32* out, u32* err) {
    HEU_LAUNCH((k_encrypt_big<L, K, W>), n, L, st, P, m, mw, r, rw, n, out, err);
  }
  static void randomize(hipStream_t st, const PubArgs& P, const u32* ct, const u32* r, int rw, long n, u32* out,
                        u32* err) {
    HEU_LAUNCH((k_randomize<L, K, W>), n, L, st, P, ct, r, rw, n, out, err);
  }
  static void mulmod(hipStream_t st, const ModArgs& mod, const u32* a, const u32* b, long n, u32* out) {
    HEU_LAUNCH((k_mulmod<L, K, W>), n, L, st, mod, a, b, n, out);
  }
  static void add_pt_i64(hipStream_t st, const PubArgs& P, const u32* ct, const i64* pt, long n, u32* out) {
    HEU_LAUNCH((k_add_pt_i64<L, K, W>), n, L, st, P, ct, pt, n, out);
  }
  static void pow_i64(hipStream_t st, const ModArgs& mod, const u32* ct, const i64* k, long n, u32* out, u8* neg,
                      u32* tab) {
    HEU_LAUNCH((k_pow_i64<L, K, W>), n, L, st, mod, ct, k, n, out, neg, tab);
  }
  static void pow_words(hipStream_t st, const ModArgs& mod, const u32* ct, const u32* kmag, const u8* kneg, int kw,
                        long n, u32* out, u8* neg, u32* tab) {
    HEU_LAUNCH((k_pow_words<L, K, W>), n, L, st, mod, ct, kmag, kneg, kw, n, out, neg, tab);
  }
  static void reduce_sum(hipStream_t st, const ModArgs& mod, bool zin, const u32* in, const Chunk* chunks, long nch,
                         u32* part, const i64* idx) {
    if (zin) HEU_LAUNCH((k_reduce_sum<L, K, W, true>), nch, L, st, mod, in, chunks, nch, part, idx);
    else HEU_LAUNCH((k_reduce_sum<L, K, W, false>), nch, L, st, mod, in, chunks, nch, part, nullptr);
  }
  static void to_mform(hipStream_t st, const ModArgs& mod, const u32* ct, long n, u32* out) {
    HEU_LAUNCH((k_to_mform<L, K, W>), n, L, st, mod, ct, n, out);
  }
  static void mulmod_m(hipStream_t st, const ModArgs& mod, const u32* a, const u32* b, long n, u32* out) {
    HEU_LAUNCH((k_mulmod_m<L, K, W>), n, L, st, mod, a, b, n, out);
  }
  static void reduce_dot0(hipStream_t st, const ModArgs& mod, const u32* in, const i64* w, const Chunk* chunks,
                          long nch, u32* pp, u32* pn, u32* tab, const i64* idx) {
    HEU_LAUNCH((k_reduce_dot0<L, K, W>), nch, L, st, mod, in, w, chunks, nch, pp, pn, tab, idx);
  }
  static void prefix_mul(hipStream_t st, const ModArgs& mod, const u32* in, const Chunk* segs, long nseg, u32* out) {
    HEU_LAUNCH((k_prefix_mul<L, K, W>), nseg, L, st, mod, in, segs, nseg, out);
  }
  static void reduce_mont(hipStream_t st, const ModArgs& mod, const u32* in, const Chunk* chunks, long nch,
                          u32* part) {
    HEU_LAUNCH((k_reduce_mont<L, K, W>), nch, L, st, mod, in, chunks, nch, part);
  }
  static void seg_finish(hipStream_t st, const ModArgs& mod, const u32* part, long nseg, u32* out) {
    HEU_LAUNCH((k_seg_finish<L, K, W>), nseg, L, st, mod, part, nseg, out);
  }
  static void mont_times_z(hipStream_t st, const ModArgs& mod, const u32* xm, const u32* yz, long n, u32* out) {
    HEU_LAUNCH((k_mont_times_z<L, K, W>), n, L, st, mod, xm, yz, n, out);
  }
  static void inv_prefix(hipStream_t st, const ModArgs& mod, bool zin, const u32* in, InvSel s, long n, int chunk,
                         u32* prefix, u32* totals) {
    long groups = (n + chunk - 1) / chunk;
    if (zin) HEU_LAUNCH((k_inv_prefix<L, K, W, true>), groups, L, st, mod, in, s, n, chunk, prefix, totals);
    else HEU_LAUNCH((k_inv_prefix<L, K, W, false>), groups, L, st, mod, in, s, n, chunk, prefix, totals);
  }
  static void inv_down(hipStream_t st, const ModArgs& mod, bool zin, const u32* in, InvSel s, long n, int chunk,
                       u32* prefix, const u32* tot_inv, u32* out) {
    long groups = (n + chunk - 1) / chunk;
    if (zin) HEU_LAUNCH((k_inv_down<L, K, W, true>), groups, L, st, mod, in, s, n, chunk, prefix, tot_inv, out);
    else HEU_LAUNCH((k_inv_down<L, K, W, false>), groups, L, st, mod, in, s, n, chunk, prefix, tot_inv, out);
  }
  // idx receives the positions of the selected elements and *nsel their count
  // (cleared first by a kernel, so the pair can be captured in a graph)
  static void inv_compact(hipStream_t st, const u8* sel, long n, u32* idx, u32* nsel) {
    hipLaunchKernelGGL(k_zero_u32<0>, dim3(1), dim3(64), 0, st, nsel);
    if (n > 0)
      hipLaunchKernelGGL(k_inv_compact<0>, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, sel, n, idx,
                         nsel);
  }
  static void inv_single(hipStream_t st, const ModArgs& mod, const u32* m_words, u32* xio, u32* err) {
    hipLaunchKernelGGL((k_inv_single<L, K, W>), dim3(1), dim3(128), 0, st, mod, m_words, xio, err);
  }
  static void mul_const(hipStream_t st, const ModArgs& mod, const u32* in, const u32* cz, long n, u32* out) {
    HEU_LAUNCH((k_mul_const<L, K, W>), n, L, st, mod, in, cz, n, out);
  }
  static void dec_half(hipStream_t st, const DecArgs& D, const u32* ct, long n, u32* halves, u32* tab) {
    if (n > 0)
      hipLaunchKernelGGL((k_dec_half<L, K, W>), dim3(blocks_for(n, L), 2), dim3(kBlock), 0, st, D, ct, n, halves, tab);
  }
  template <bool kDec>
  static void worklist(hipStream_t st, const WorkItem* items, int nitems, u32 total_units, u32* ctr, u32* slot_tab,
                       unsigned grid) {
    if (total_units > 0 && grid > 0)
      hipLaunchKernelGGL((k_worklist<L, K, W, kDec>), dim3(grid), dim3(kBlock), 0, st, items, nitems, total_units, ctr,
                         slot_tab);
  }
  template <bool kDec>
  static int resident() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_worklist<L, K, W, kDec>, kBlock, 0) != hipSuccess) return 1;
    return n > 0 ? n : 1;
  }
  static DecShape dec_shape(const WorklistShape& wl) {
    DecShape d;
    d.L = L;
    d.K = K;
    d.dec_half = dec_half;
    d.wl = wl;
    return d;
  }
  template <bool kDec>
  static WorklistShape worklist_shape() {
    WorklistShape w;
    w.launch = worklist<kDec>;
    w.resident = resident<kDec>;
    w.groups_per_unit = kBlock / L;
    w.slot_table_words = (long)(L * K) << (kDec ? kDecWin<W> - 1 : kPowWin);
    return w;
  }
};

template <int PW>
static void launch_combine(hipStream_t st, const CombArgs& C, const u32* halves, long n, u32* out, int ow, u8* ovf) {
  if (n > 0)
    hipLaunchKernelGGL((k_dec_combine<PW>), dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, C,
                       halves, n, out, ow, ovf);
}

// Shapes (29-bit limbs, 29*NL >= bits(m) + 2):
//   2048-bit keys: n^2 over 4 lanes x 36 limbs, p^2 over 2 lanes x 36 limbs
//                  (or 4 lanes x 18);
//   3072-bit keys: n^2 over 8 lanes x 27 limbs, p^2 over 4 lanes x 27 limbs
//                  (or 8 lanes x 14: NL = 112 >= 106).
// W = words of the ciphertext (decryption halves hold W/2).
typedef Shape<4, 36, 128> ShapeN2048;
typedef Shape<8, 27, 192> ShapeN3072;
typedef Shape<2, 36, 128> ShapeP2048;
typedef Shape<4, 18, 128> ShapeP2048w;
typedef Shape<4, 27, 192> ShapeP3072;
typedef Shape<8, 14, 192> ShapeP3072w;

// The launch table of one key size; the work-list kernels come from their own
// translation units (they are the largest kernels to compile).
template <class SN, int PW>
static Ops make_ops(const WorklistShape& wl_n, const DecShape& dec0, const DecShape& dec1) {
  Ops o;
  o.n_L = SN::kL;
  o.n_K = SN::kK;
  o.ct_words = SN::kW;
  o.p_words = PW;
  o.table_level = Launch<SN>::table_level;
  o.encrypt_i64 = Launch<SN>::encrypt_i64;
  o.encrypt_big = Launch<SN>::encrypt_big;
  o.randomize = Launch<SN>::randomize;
  o.mulmod = Launch<SN>::mulmod;
  o.to_mform = Launch<SN>::to_mform;
  o.mulmod_m = Launch<SN>::mulmod_m;
  o.add_pt_i64 = Launch<SN>::add_pt_i64;
  o.pow_i64 = Launch<SN>::pow_i64;
  o.pow_words = Launch<SN>::pow_words;
  o.reduce_sum = Launch<SN>::reduce_sum;
  o.reduce_dot0 = Launch<SN>::reduce_dot0;
  o.reduce_mont = Launch<SN>::reduce_mont;
  o.seg_finish = Launch<SN>::seg_finish;
  o.mont_times_z = Launch<SN>::mont_times_z;
  o.inv_prefix = Launch<SN>::inv_prefix;
  o.inv_down = Launch<SN>::inv_down;
  o.inv_compact = Launch<SN>::inv_compact;
  o.inv_single = Launch<SN>::inv_single;
  o.mul_const = Launch<SN>::mul_const;
  o.prefix_mul = Launch<SN>::prefix_mul;
  o.combine = launch_combine<PW>;
  o.wl_n = wl_n;
  o.dec[0] = dec0;
  o.dec[1] = dec1;
  return o;
}

}  // namespace heu_hip
