// kernels.hpp -- argument blocks shared by the host engine and the gfx950
// kernels, and the per-key-size table of launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace heu_hip {

typedef uint8_t u8;

// Device error flags (OR-ed by kernels, read by the engine at sync).
constexpr uint32_t kErrRange = 1u;          // an input outside the supported range
constexpr uint32_t kErrNotInvertible = 2u;  // gcd(x, n^2) != 1 in an inversion

// One Montgomery modulus in R29 form (29-bit limbs, interleaved for its group size).
struct ModArgs {
  const uint32_t* m;    // NL words
  const uint32_t* one;  // R mod m
  const uint32_t* r2;   // R^2 mod m
  const uint32_t* r3;   // R^3 mod m (p^2 and q^2 only)
  uint32_t mprime;      // -m^-1 mod 2^29
  // n^2 only: entry k (NL words, k = 0 .. kRpowEntries-1) is R^(k+1) mod m,
  // the Montgomery form of R^k.  A Z-space chunk of k terms multiplied by its
  // entry leaves the chunk's product in Montgomery form.
  const uint32_t* rpow;
};
constexpr int kRpowEntries = 33;  // chunks of up to 32 Z-space terms

// Public-key side, modulus n^2.
struct PubArgs {
  ModArgs n2;
  const uint32_t* n29;    // n as plain 29-bit limbs (NL words)
  const uint32_t* table;  // fixed-base table of h_s: windows x 2^unit entries x NL words
  int windows, unit;
};

// Secret-key side: the two CRT halves mod p^2 and q^2.
struct DecArgs {
  ModArgs p2, q2;
  const uint32_t* sched_p;  // sliding-window schedule of p-1
  const uint32_t* sched_q;  // ... and of q-1
  int nsched_p, nsched_q;
};

// CRT recombination constants, 32-bit words (PW = words of p).
struct CombArgs {
  const uint32_t *p, *q;          // PW words
  const uint32_t *pinv, *qinv;    // p^-1, q^-1 mod 2^(32 PW)
  const uint32_t *hpR, *hqR;      // hp * 2^(32 PW) mod p, hq * 2^(32 PW) mod q
  const uint32_t* pinvqR;         // (p^-1 mod q) * 2^(32 PW) mod q
  const uint32_t *n, *n_half;     // 2 PW words
  uint32_t pprime, qprime;        // -p^-1, -q^-1 mod 2^32
};

// One contiguous run of a reduction segment.
struct Chunk {
  long start, len;
};

// Work-list launches: one persistent launch serves many items (vectors of any
// length, of any key of the launch's shape).  Units are blocks of groups; a
// workgroup takes the next unit from an atomic counter, finds its item by
// binary search over unit0 and runs the op's per-element body.
enum WorkOp : int {
  kWorkEncrypt = 0,  // a: int64 m, b: r (aux words each), out: ct
  kWorkRandomize,    // a: ct or null (EncryptZero), b: r (aux words), out: ct
  kWorkAdd,          // a, b: ct, out: ct
  kWorkAddPt,        // a: ct, b: int64 pt, out: ct
  kWorkMul,          // a: ct, b: int64 scalar, out: ct (|k| power), neg: sign flags
  kWorkDot0,         // a: ct, b: int64 weights, c: chunks, d: index list or null, out/out2: pos/neg partials
  kWorkSum0,         // a: ct, c: chunks, d: index list or null, out: partials (count = chunks)
  kWorkDecHalf       // a: ct, c: window schedule (aux ops), out: this half's outputs; pub.n2 = p^2 or q^2
};

struct WorkItem {
  long unit0;  // first unit of this item in the launch (items sorted by unit0)
  long count;  // elements, or chunks for the reductions
  int op, aux;
  PubArgs pub;  // pub.n2 is the item's modulus
  const void* a;
  const void* b;
  const void* c;
  const void* d;  // kWorkDot0 / kWorkSum0: optional int64 index list (gathered sums)
  void* out;
  void* out2;
  u8* neg;
  uint32_t* err;
};

typedef void (*TableLevelFn)(hipStream_t, const ModArgs&, const uint32_t*, uint32_t*, int, int, int);
typedef void (*EncI64Fn)(hipStream_t, const PubArgs&, const int64_t*, const uint32_t*, int, long, uint32_t*,
                         uint32_t*);
typedef void (*EncBigFn)(hipStream_t, const PubArgs&, const uint32_t*, int, const uint32_t*, int, long, uint32_t*,
                         uint32_t*);
typedef void (*RandFn)(hipStream_t, const PubArgs&, const uint32_t*, const uint32_t*, int, long, uint32_t*,
                       uint32_t*);
typedef void (*MulModFn)(hipStream_t, const ModArgs&, const uint32_t*, const uint32_t*, long, uint32_t*);
typedef void (*ConvFn)(hipStream_t, const ModArgs&, const uint32_t*, long, uint32_t*);
typedef void (*AddPtFn)(hipStream_t, const PubArgs&, const uint32_t*, const int64_t*, long, uint32_t*);
typedef void (*PowFn)(hipStream_t, const ModArgs&, const uint32_t*, const int64_t*, long, uint32_t*, u8*, uint32_t*);
typedef void (*PowWordsFn)(hipStream_t, const ModArgs&, const uint32_t*, const uint32_t*, const u8*, int, long,
                           uint32_t*, u8*, uint32_t*);
// zin: Z-space words at level 0 (else R29 Montgomery limbs); either way each
// partial is the Montgomery form of its chunk's product (an empty chunk: of 1)
typedef void (*ReduceSumFn)(hipStream_t, const ModArgs&, bool, const uint32_t*, const Chunk*, long, uint32_t*,
                            const int64_t*);
typedef void (*ReduceDot0Fn)(hipStream_t, const ModArgs&, const uint32_t*, const int64_t*, const Chunk*, long,
                             uint32_t*, uint32_t*, uint32_t*, const int64_t*);
typedef void (*PrefixMulFn)(hipStream_t, const ModArgs&, const uint32_t*, const Chunk*, long, uint32_t*);
typedef void (*ReduceMontFn)(hipStream_t, const ModArgs&, const uint32_t*, const Chunk*, long, uint32_t*);
// Montgomery-form values -> canonical Z-space words
typedef void (*SegFinishFn)(hipStream_t, const ModArgs&, const uint32_t*, long, uint32_t*);
typedef void (*MontTimesZFn)(hipStream_t, const ModArgs&, const uint32_t*, const uint32_t*, long, uint32_t*);
// Which level-0 elements a batch inversion works on: all of them (both null),
// those with sel[i] != 0, or positions [0, *nsel) of the index list idx.
struct InvSel {
  const u8* sel;
  const uint32_t* idx;
  const uint32_t* nsel;
};
typedef void (*InvPrefixFn)(hipStream_t, const ModArgs&, bool, const uint32_t*, InvSel, long, int, uint32_t*,
                            uint32_t*);
typedef void (*InvDownFn)(hipStream_t, const ModArgs&, bool, const uint32_t*, InvSel, long, int, uint32_t*,
                          const uint32_t*, uint32_t*);
typedef void (*InvCompactFn)(hipStream_t, const u8*, long, uint32_t*, uint32_t*);
typedef void (*InvSingleFn)(hipStream_t, const ModArgs&, const uint32_t*, uint32_t*, uint32_t*);
typedef void (*MulConstFn)(hipStream_t, const ModArgs&, const uint32_t*, const uint32_t*, long, uint32_t*);
typedef void (*DecHalfFn)(hipStream_t, const DecArgs&, const uint32_t*, long, uint32_t*, uint32_t*);
typedef void (*CombineFn)(hipStream_t, const CombArgs&, const uint32_t*, long, uint32_t*, int, u8*);
// items (device), item count, total units, counters (2 words, zero between
// launches: the kernel's last workgroup clears them), per-slot window tables,
// grid size (workgroups)
typedef void (*WorklistFn)(hipStream_t, const WorkItem*, int, uint32_t, uint32_t*, uint32_t*, unsigned);
// Resident workgroups per CU of a work-list kernel (occupancy query).
typedef int (*ResidentFn)();

// Geometry of one work-list kernel.
struct WorklistShape {
  WorklistFn launch;
  ResidentFn resident;
  int groups_per_unit;      // elements per unit (kBlock / L)
  long slot_table_words;    // window-table words per group slot
};

// One decryption shape: the lane group mod p^2 / q^2 (L lanes x K limbs) and
// its kernels.  Each key size has kDecShapes of them (lanes per bignum is a
// tuning choice: fewer lanes, more throughput per lane; more lanes, shorter
// latency per element).
struct DecShape {
  int L, K;
  DecHalfFn dec_half;
  WorklistShape wl;
};
constexpr int kDecShapes = 2;

// Launchers and shape constants for one key size.
struct Ops {
  int n_L, n_K, ct_words;  // group shape mod n^2 and ciphertext words
  int p_words;             // words of p
  TableLevelFn table_level;
  EncI64Fn encrypt_i64;
  EncBigFn encrypt_big;
  RandFn randomize;
  MulModFn mulmod;
  ConvFn to_mform;    // Z-space words -> engine form (k_to_mform)
  MulModFn mulmod_m;  // engine-form ct (+) ct
  AddPtFn add_pt_i64;
  PowFn pow_i64;
  PowWordsFn pow_words;
  ReduceSumFn reduce_sum;
  ReduceDot0Fn reduce_dot0;
  ReduceMontFn reduce_mont;
  SegFinishFn seg_finish;
  MontTimesZFn mont_times_z;
  InvPrefixFn inv_prefix;
  InvDownFn inv_down;
  InvCompactFn inv_compact;
  InvSingleFn inv_single;
  MulConstFn mul_const;
  PrefixMulFn prefix_mul;
  CombineFn combine;
  WorklistShape wl_n;             // n^2 ops: encrypt, randomize, add, add_pt, mul, dot0, sum0
  DecShape dec[kDecShapes];       // CRT decryption halves mod p^2 / q^2; [0] is the default
};

const Ops* ops_for_key_bits(int key_bits);

constexpr int kBlock = 256;         // threads per workgroup of every group kernel
constexpr int kPowWindowBits = 4;  // fixed window of k_pow_i64 / dot terms
// Sliding window of the CRT exponentiations, per key size.  3072-bit keys use 6
// bits: their 1536-bit exponents take about 251 table and window products
// instead of 272 at 5 bits, and decryption ran 1.2% faster.  2048-bit keys keep
// 5 bits: there the 6-bit window saves 9 products, but its tables are twice as
// large, and the two came out even.
constexpr int dec_window_bits(int key_bits) { return key_bits >= 3072 ? 6 : 5; }

}  // namespace heu_hip
