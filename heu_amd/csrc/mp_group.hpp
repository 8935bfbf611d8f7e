// mp_group.hpp -- multi-precision core for gfx950: one big residue spread over
// a group of L lanes of a 64-wide wavefront.
//
// Representation ("R29"): a residue mod an odd m is NL = L*K limbs of 29 bits.
// Lane l of the group owns limbs [l*K, l*K + K) ("blocked" ownership, so the
// one-limb shift of Montgomery's CIOS reduction moves one value per lane).
// In global memory a residue is stored "interleaved": word t*L + l holds limb
// l*K + t, so the L lanes of a group load L consecutive dwords per step.
//
// Montgomery multiplication is CIOS over radix 2^29 with 64-bit lazy
// accumulators.  gfx950 issues v_mad_u64_u32 at the full VALU rate, the same
// as an add, so the loop is built to spend as few non-multiply instructions as
// possible per product: every 29x29-bit product is one v_mad_u64_u32 into its
// accumulator with no carry chain.  An accumulator gains < 2^59.01 per outer
// iteration, and lives in one lane for at most K iterations: it enters as the
// lane's top slot holding < 2^29, and leaves as slot 0, whose bits above 29 go
// to slot 1 of the same lane and whose low 29 bits go to the lane below.  So
// for K <= 31 no carry step is needed at all.  K = 36 products (squarings'
// bound is 27 iterations, below) carry one slot per iteration, the one in the
// middle of the lane, which hands its bits above 32 to the slot above it (one
// multiply-add by 2^3) and so cuts every slot's life into runs of <= K/2 + 1
// iterations between carries.  With R = 2^(29*NL) >= 4m, inputs below 2m give
// outputs below 2m, so chained multiplications never need the conditional
// subtraction; only results leaving the engine are canonicalised.
//
// Cross-lane traffic per outer iteration: one DPP broadcast of the Montgomery
// quotient q from group lane 0 and one 32-bit DPP shift of the limb leaving
// each lane.  The multiplier limb a_i comes from an LDS copy of the operand,
// read by all lanes of the group at the same address (a broadcast read).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace heu_hip {

typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

constexpr int kLimbBits = 29;
constexpr u32 kLimbMask = (1u << kLimbBits) - 1;
constexpr int kNormEvery = 31;  // longest run of multiply iterations without a carry (see bound above)

// ---------------------------------------------------------------------------
// DPP controls (GFX9 encoding) and group primitives
// ---------------------------------------------------------------------------
constexpr int kDppRowShl1 = 0x101;      // lane i <- lane i+1 within a row of 16 (lane 15 <- 0)
constexpr int kDppRowShr1 = 0x111;      // lane i <- lane i-1 within a row of 16 (lane 0 <- 0)
constexpr int kDppRowNewBcast = 0x150;  // + n: every lane of the row <- lane n of the row
constexpr int quad_perm(int a, int b, int c, int d) { return a | (b << 2) | (c << 4) | (d << 6); }

template <int L>
__device__ __forceinline__ u32 grp_lane() {
  return threadIdx.x & (u32)(L - 1);
}

// Value of group lane 0, in every lane of the group.
template <int L>
__device__ __forceinline__ u32 grp_bcast0(u32 x) {
  static_assert(L == 2 || L == 4 || L == 8 || L == 16, "group size");
  if constexpr (L == 16) {
    return __builtin_amdgcn_update_dpp(x, x, kDppRowNewBcast + 0, 0xf, 0xf, false);
  } else if constexpr (L == 8) {
    u32 v = __builtin_amdgcn_update_dpp(x, x, kDppRowNewBcast + 0, 0xf, 0x3, false);
    return __builtin_amdgcn_update_dpp(v, x, kDppRowNewBcast + 8, 0xf, 0xc, false);
  } else if constexpr (L == 4) {
    return __builtin_amdgcn_update_dpp(x, x, quad_perm(0, 0, 0, 0), 0xf, 0xf, false);
  } else {
    return __builtin_amdgcn_update_dpp(x, x, quad_perm(0, 0, 2, 2), 0xf, 0xf, false);
  }
}

// Value of group lane L-1, in every lane of the group.
template <int L>
__device__ __forceinline__ u32 grp_bcast_top(u32 x) {
  if constexpr (L == 16) {
    return __builtin_amdgcn_update_dpp(x, x, kDppRowNewBcast + 15, 0xf, 0xf, false);
  } else if constexpr (L == 8) {
    u32 v = __builtin_amdgcn_update_dpp(x, x, kDppRowNewBcast + 7, 0xf, 0x3, false);
    return __builtin_amdgcn_update_dpp(v, x, kDppRowNewBcast + 15, 0xf, 0xc, false);
  } else if constexpr (L == 4) {
    return __builtin_amdgcn_update_dpp(x, x, quad_perm(3, 3, 3, 3), 0xf, 0xf, false);
  } else {
    return __builtin_amdgcn_update_dpp(x, x, quad_perm(1, 1, 3, 3), 0xf, 0xf, false);
  }
}

// Lane i receives lane i+1's value; the top lane of the group receives 0.
// (Groups of one wave may diverge, so the next group's lane is not read.)
template <int L>
__device__ __forceinline__ u32 grp_shl1(u32 x, u32 lid) {
  u32 v = __builtin_amdgcn_update_dpp(0u, x, kDppRowShl1, 0xf, 0xf, true);
  if constexpr (L < 16) v = (lid == (u32)(L - 1)) ? 0u : v;
  return v;
}

// Lane i receives lane i-1's value; group lane 0 receives 0.
template <int L>
__device__ __forceinline__ u32 grp_shr1(u32 x, u32 lid) {
  u32 v = __builtin_amdgcn_update_dpp(0u, x, kDppRowShr1, 0xf, 0xf, true);
  if constexpr (L < 16) v = (lid == 0u) ? 0u : v;
  return v;
}

// Lane i receives lane i+1's value and the top lane receives group lane 0's.
// For groups within a quad (L <= 4) this is one rotation inside the group;
// Montgomery's limb shift uses it where group lane 0 holds 0, which then
// arrives in the top lane with no select.  Wider groups shift (top lane <- 0).
template <int L>
__device__ __forceinline__ u32 grp_shl1_z(u32 x, u32 lid) {
  if constexpr (L == 2) {
    return __builtin_amdgcn_update_dpp(x, x, quad_perm(1, 0, 3, 2), 0xf, 0xf, false);
  } else if constexpr (L == 4) {
    return __builtin_amdgcn_update_dpp(x, x, quad_perm(1, 2, 3, 0), 0xf, 0xf, false);
  } else {
    return grp_shl1<L>(x, lid);
  }
}

// Groups within a quad (L <= 4): broadcast of group lane 0 and the rotation of
// grp_shl1_z as DPP moves with bound_ctrl set, which changes nothing here
// (every lane reads an active lane of its own quad) but lets an AND applied to
// the result fuse with the move into one v_and_b32 with a DPP operand.
template <int L>
__device__ __forceinline__ u32 quad_bcast0(u32 x) {
  static_assert(L == 2 || L == 4, "groups within a quad");
  return __builtin_amdgcn_mov_dpp(x, L == 2 ? quad_perm(0, 0, 2, 2) : quad_perm(0, 0, 0, 0), 0xf, 0xf, true);
}
template <int L>
__device__ __forceinline__ u32 quad_rot1(u32 x) {
  static_assert(L == 2 || L == 4, "groups within a quad");
  return __builtin_amdgcn_mov_dpp(x, L == 2 ? quad_perm(1, 0, 3, 2) : quad_perm(1, 2, 3, 0), 0xf, 0xf, true);
}

// Wider groups (L >= 8): lane i receives lane i+1's value AND mask, the top
// lane of the group 0.  The DPP shift (bound_ctrl: lane 15 of a row reads 0)
// fuses with the AND; the top lane of a group inside the row is cleared by a
// select, as the next group's lane is not read (groups may diverge).
__device__ __forceinline__ u32 row_shl1_masked(u32 x, u32 mask, bool top) {
  const u32 v = __builtin_amdgcn_mov_dpp(x, kDppRowShl1, 0xf, 0xf, true) & mask;
  return top ? 0u : v;
}

template <int L>
__device__ __forceinline__ u64 grp_shr1_64(u64 x, u32 lid) {
  u32 lo = grp_shr1<L>((u32)x, lid), hi = grp_shr1<L>((u32)(x >> 32), lid);
  return ((u64)hi << 32) | lo;
}

// Orders this wave's LDS traffic around a hand-off between lanes: a compiler
// barrier plus wavefront-scope fences (LDS is in-order within a wave).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ u64 mad64(u32 a, u32 b, u64 c) { return (u64)a * b + c; }

// Any active lane of the wave has p set (wave-uniform).
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// ---------------------------------------------------------------------------
// Compile-time loop helper
// ---------------------------------------------------------------------------
template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// ---------------------------------------------------------------------------
// Modulus of a group: its limbs (interleaved R29 layout, staged in LDS once per
// block and shared by all groups of the block) and -m^-1 mod 2^29.
// ---------------------------------------------------------------------------
template <int L, int K>
struct GMod {
  const u32* m;  // LDS, word t*L + lane holds limb lane*K + t
  u32 mprime;
};

// Interleaved global layout <-> lane registers.
template <int L, int K>
__device__ __forceinline__ void load_r29(u32 (&x)[K], const u32* __restrict__ src, u32 lid) {
#pragma unroll
  for (int t = 0; t < K; ++t) x[t] = src[t * L + lid];
}
template <int L, int K>
__device__ __forceinline__ void store_r29(u32* __restrict__ dst, const u32 (&x)[K], u32 lid) {
#pragma unroll
  for (int t = 0; t < K; ++t) dst[t * L + lid] = x[t];
}

// The group's multiplier operand in LDS, plain limb order (limb j at a[j]).
template <int L, int K>
__device__ __forceinline__ void put_lds(u32* __restrict__ a_lds, const u32 (&x)[K], u32 lid) {
  wave_lds_sync();
#pragma unroll
  for (int t = 0; t < K; ++t) a_lds[lid * K + t] = x[t];
  wave_lds_sync();
}
template <int L, int K>
__device__ __forceinline__ void get_lds(u32 (&x)[K], const u32* __restrict__ a_lds, u32 lid) {
  wave_lds_sync();
#pragma unroll
  for (int t = 0; t < K; ++t) x[t] = a_lds[lid * K + t];
}

// acc = a * b * R^-1 (mod m), lazy: result < 2m when a, b < 2m.
// a: the group's LDS operand (NL limbs, each < 2^29 + 65);
// b: this lane's K limbs (each < 2^29 + 65).
//
// Per outer iteration i (a_i = limb i of a):
//   acc += a_i * b;  q = acc_0 * m' mod 2^29 (group lane 0, broadcast);
//   acc += q * m;    acc_1 += acc_0 >> 29 in every lane;
//   the low 29 bits of acc_0 move to lane-1's top slot.
// In group lane 0, acc_0 = 0 mod 2^29, so nothing is lost there (and the zero
// it sends is what the top lane receives); in other lanes the split keeps the
// value: acc_0 = lo + 2^29 * hi.
//
// Squaring (kSqr, b = the lane's limbs of the LDS operand a): at iteration i,
// lane l multiplies a_i by its limbs a_{lK+j}, so the full product forms every
// cross term a_x a_y twice, at iteration x and at iteration y.  Which products
// an unrolled iteration issues must not depend on the lane, so the choice is a
// function of (i mod K, j) only:
//   j == i mod K: a_i * a_j as in the full product (the square a_i^2 in one
//     lane; cross terms of equal residues, formed at both iterations);
//   otherwise: 2 a_i * a_j if sq_cross_here<K>(i mod K, j), else nothing.
// sq_cross_here is a tournament on the K residues (exactly one of (x, y),
// (y, x) holds), so every other cross term is formed once, doubled.  An
// iteration then issues (K+1)/2 product multiply-adds instead of K.  Every
// position's products are all added by the iteration that consumes it, so the
// quotients, and the result, equal the full product's.
// Bound: in units of 2^58, an accumulator gains < 1 per iteration from q*m and
// 2 (doubled), 1 (j == i mod K) or 0 from the product.  Enumerating the
// schedule, any G consecutive iterations of one slot in one lane total at most
// G + P(G) units with G + P(G) <= 63 for G <= 27 (P(G) = the largest sum of
// product weights over such a window), so slots stay below 2^64 in runs of up
// to 27 iterations between carries, and K <= 27 needs none.
constexpr int kNormEverySq = 27;

template <int K>
constexpr bool sq_cross_here(int x, int y) {
  const int d = (y - x + K) % K;
  if constexpr (K & 1) return d >= 1 && d <= (K - 1) / 2;
  return (d >= 1 && d < K / 2) || (d == K / 2 && x < K / 2);
}

template <int L, int K, bool kSqr>
__device__ __forceinline__ void mont_core(u64 (&acc)[K], const u32* __restrict__ a_lds, const u32 (&b)[K],
                                          const GMod<L, K>& md) {
  constexpr int kNorm = kSqr ? kNormEverySq : kNormEvery;
  // Carries when K > kNorm: the slot in the middle of the lane, once per
  // iteration (runs of K-1-kMid and kMid+1 iterations).  In a multiplication a
  // scheduling barrier ends each iteration: otherwise the compiler interleaves
  // iterations until the K = 36 kernels spill (k_dec_half: 724 B of scratch).
  // With it they spill less than a full carry step over all slots every kNorm
  // iterations made them spill, and issue fewer instructions.
  constexpr int kMid = K / 2;
  constexpr bool kSlot = K > kNorm;
  static_assert(!kSlot || (K - 1 - kMid <= kNorm && kMid + 1 <= kNorm), "carry runs too long");
  // 2^(32-29) for the carry, hidden from the optimizer, which would otherwise
  // turn the one multiply-add into a 64-bit shift, two masks and an add.
  u32 hi_weight = 1u << (32 - kLimbBits);
  if constexpr (kSlot) asm volatile("" : "+s"(hi_weight));
  // The limb mask in a register: a DPP operand's partner cannot be a literal.
  u32 mask_v = kLimbMask;
  asm volatile("" : "+v"(mask_v));
#pragma unroll
  for (int t = 0; t < K; ++t) acc[t] = 0;
  const u32 lid = grp_lane<L>();
  for (int blk = 0; blk < L; ++blk) {
    const u32* ap = a_lds + blk * K;
    u32 m[K];
#pragma unroll
    for (int t = 0; t < K; ++t) m[t] = md.m[t * L + lid];
    static_for<K>([&](auto S) {
      constexpr int s = decltype(S)::value;
      const u32 ai = ap[s];
      if constexpr (kSqr) {
        const u32 ai2 = ai << 1;
        static_for<K>([&](auto J) {
          constexpr int j = decltype(J)::value;
          if constexpr (j == s)
            acc[(j + s) % K] = mad64(ai, b[j], acc[(j + s) % K]);
          else if constexpr (sq_cross_here<K>(s, j))
            acc[(j + s) % K] = mad64(ai2, b[j], acc[(j + s) % K]);
        });
      } else {
#pragma unroll
        for (int j = 0; j < K; ++j) acc[(j + s) % K] = mad64(ai, b[j], acc[(j + s) % K]);
      }
      u32 q;
      if constexpr (L <= 4)
        q = quad_bcast0<L>((u32)acc[s % K] * md.mprime) & mask_v;
      else
        q = grp_bcast0<L>(((u32)acc[s % K] * md.mprime) & kLimbMask);
#pragma unroll
      for (int j = 0; j < K; ++j) acc[(j + s) % K] = mad64(q, m[j], acc[(j + s) % K]);
      const u64 out0 = acc[s % K];
      acc[(s + 1) % K] += out0 >> kLimbBits;
      // group lane 0 sends 0 (its low 29 bits are 0; bits 29-31 are masked off)
      if constexpr (L <= 4)
        acc[s % K] = (u64)(quad_rot1<L>((u32)out0) & mask_v);
      else
        acc[s % K] = (u64)(row_shl1_masked((u32)out0, mask_v, lid == (u32)(L - 1)));
      if constexpr (kSlot) {
        // The slot that is logical kMid in the next iteration, split at bit 32:
        // its high word is worth 2^3 units of the slot above (one multiply-add).
        // The low word left behind (< 2^32) is as harmless as a 29-bit limb.
        u64& x = acc[(kMid + s + 1) % K];
        u64& above = acc[(kMid + 1 + s + 1) % K];
        above = mad64((u32)(x >> 32), hi_weight, above);
        x = (u32)x;
      }
      if constexpr (kSlot && !kSqr) __builtin_amdgcn_sched_barrier(0);
    });
  }
}

template <int L, int K>
__device__ __forceinline__ void mont_mul(u64 (&acc)[K], const u32* __restrict__ a_lds, const u32 (&b)[K],
                                         const GMod<L, K>& md) {
  mont_core<L, K, false>(acc, a_lds, b, md);
}

// acc = a * a * R^-1 (mod m), lazy; b holds this lane's limbs of a_lds.
template <int L, int K>
__device__ __forceinline__ void mont_sqr(u64 (&acc)[K], const u32* __restrict__ a_lds, const u32 (&b)[K],
                                         const GMod<L, K>& md) {
  mont_core<L, K, true>(acc, a_lds, b, md);
}

// Lazy accumulators -> limbs each < 2^29 + 65 (value unchanged; value < 2^(29*NL)).
template <int L, int K>
__device__ __forceinline__ void seminorm(const u64 (&acc)[K], u32 (&out)[K]) {
  const u32 lid = grp_lane<L>();
  u64 c = 0;
#pragma unroll
  for (int t = 0; t < K; ++t) {
    u64 v = acc[t] + c;
    out[t] = (u32)v & kLimbMask;
    c = v >> kLimbBits;
  }
  u64 cin = grp_shr1_64<L>(c, lid);
  u64 v = (u64)out[0] + cin;
  out[0] = (u32)v & kLimbMask;
  out[1] += (u32)(v >> kLimbBits);
}

// Signed lazy limbs -> canonical limbs in [0, 2^29).  The carry discarded
// above the top limb (a signed count of 2^(29*NL)) is returned in the top
// lane; it is 0 whenever the value lies in [0, 2^(29*NL)).
template <int L, int K>
__device__ __forceinline__ i64 norm_full(const i64 (&v)[K], u32 (&out)[K]) {
  const u32 lid = grp_lane<L>();
  i64 c = 0;
#pragma unroll
  for (int t = 0; t < K; ++t) {
    i64 x = v[t] + c;
    out[t] = (u32)x & kLimbMask;
    c = x >> kLimbBits;
  }
  i64 dropped = 0;
  for (;;) {
    if (lid == (u32)(L - 1)) {
      dropped += c;
      c = 0;
    }
    i64 cin = (i64)grp_shr1_64<L>((u64)c, lid);
    if (!wave_any(cin != 0)) break;
    c = cin;
#pragma unroll
    for (int t = 0; t < K; ++t) {
      i64 x = (i64)out[t] + c;
      out[t] = (u32)x & kLimbMask;
      c = x >> kLimbBits;
    }
  }
  return dropped;
}

// Value < 2m (limbs < 2^29 + 65) -> canonical residue in [0, m).
template <int L, int K>
__device__ __forceinline__ void canonicalize(u32 (&x)[K], const GMod<L, K>& md) {
  i64 v[K];
#pragma unroll
  for (int t = 0; t < K; ++t) v[t] = (i64)x[t];
  u32 u[K];
  norm_full<L, K>(v, u);
#pragma unroll
  for (int t = 0; t < K; ++t) v[t] = (i64)u[t] - (i64)md.m[t * L + grp_lane<L>()];
  u32 d[K];
  i64 drop = norm_full<L, K>(v, d);
  u32 neg = grp_bcast_top<L>(drop < 0 ? 1u : 0u);
#pragma unroll
  for (int t = 0; t < K; ++t) x[t] = neg ? u[t] : d[t];
}

// 32-bit little-endian words staged in LDS -> this lane's limbs starting at
// limb `first` (limbs past the number are 0).
// kFresh hides the lane index from the optimizer, so that the K word offsets
// are computed here, at each use, instead of being hoisted out of a caller's
// loop, where they would stay live across the Montgomery products and spill.
template <int L, int K, bool kFresh = false>
__device__ __forceinline__ void words_to_limbs(const u32* __restrict__ w_lds, int words, int first, u32 (&x)[K]) {
  u32 lid = grp_lane<L>();
  if constexpr (kFresh) asm volatile("" : "+v"(lid));
#pragma unroll
  for (int t = 0; t < K; ++t) {
    int j = first + (int)lid * K + t;
    int bit = kLimbBits * j, w0 = bit >> 5, s = bit & 31;
    u64 lo = w0 < words ? w_lds[w0] : 0u;
    u64 hi = w0 + 1 < words ? w_lds[w0 + 1] : 0u;
    x[t] = (u32)(((hi << 32) | lo) >> s) & kLimbMask;
  }
}

// Canonical limbs (plain order in LDS) -> output word w.
template <int L, int K>
__device__ __forceinline__ u32 limbs_to_word(const u32* __restrict__ l_lds, int w) {
  constexpr int NL = L * K;
  int bit = 32 * w, j0 = bit / kLimbBits, s = bit - kLimbBits * j0;
  u64 v = j0 < NL ? l_lds[j0] : 0u;
  if (j0 + 1 < NL) v |= (u64)l_lds[j0 + 1] << kLimbBits;
  if (j0 + 2 < NL) v |= (u64)l_lds[j0 + 2] << (2 * kLimbBits);
  return (u32)(v >> s);
}

// Load a W-word little-endian number from global memory into the group's LDS
// staging area (coalesced: lane l loads words l, l+L, ...).
template <int L>
__device__ __forceinline__ void stage_words(u32* __restrict__ w_lds, const u32* __restrict__ src, int words) {
  const u32 lid = grp_lane<L>();
  wave_lds_sync();
  for (int w = (int)lid; w < words; w += L) w_lds[w] = src[w];
  wave_lds_sync();
}

// Store canonical limbs as `words` little-endian 32-bit words.
template <int L, int K>
__device__ __forceinline__ void store_words(u32* __restrict__ dst, u32* __restrict__ l_lds, const u32 (&x)[K],
                                            int words) {
  const u32 lid = grp_lane<L>();
  put_lds<L, K>(l_lds, x, lid);
  for (int w = (int)lid; w < words; w += L) dst[w] = limbs_to_word<L, K>(l_lds, w);
}

}  // namespace heu_hip
