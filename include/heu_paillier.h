/*
 * heu_paillier.h -- C ABI of the MI355X Paillier batch engine.
 *
 * This replaces the C ABI that HEU's CUDA backend binds,
 * heu/library/algorithms/paillier_gpu/gpulib/gpupaillier.h:17-91 (one
 * synchronous `int gpu_paillier_<op>(...)` per op over 512-byte host structs,
 * device 0 only, allocation per call).  Differences by design:
 *   - a context (device + stream) and resident keys are explicit handles; keys
 *     keep n, n^2, Montgomery constants, CRT constants and the fixed-base table
 *     of h_s on the device;
 *   - batch entry points take DEVICE pointers to flat arrays of little-endian
 *     uint32 limbs and are asynchronous on the context's stream;
 *     heu_ctx_sync() waits and reports device-side errors;
 *   - ciphertexts are the Z-space value c = (1 + n*m) * h_s^r mod n^2 of
 *     heu/library/algorithms/paillier_zahlen/encryptor.cc:42-49, stored as
 *     ct_words = 2*n_words limbs (128 for 2048-bit keys, 192 for 3072-bit).
 *     heu_to_mont()/heu_from_mont() convert to the Montgomery form that
 *     paillier_z::Ciphertext::c_ stores, for a stated word size.
 *   - every entry point returns a status; heu_last_error() gives the message.
 *
 * Signed plaintexts cross as two's complement over a stated number of limbs.
 * Empty work is a successful no-op: count = 0 (or no segments, or segments
 * with no terms) reads and writes nothing, and pointers to empty arrays may
 * be NULL.
 */
#ifndef HEU_PAILLIER_H_
#define HEU_PAILLIER_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct heu_ctx heu_ctx;
typedef struct heu_key heu_key;

enum {
  HEU_OK = 0,
  HEU_EINVAL = 1,       /* bad argument */
  HEU_ERANGE = 2,       /* plaintext / randomness / output out of range */
  HEU_ENOTINV = 3,      /* a value has no inverse mod n^2 */
  HEU_ENOMEM = 4,       /* device allocation failed */
  HEU_EDEVICE = 5,      /* HIP runtime error */
  HEU_EUNSUPPORTED = 6, /* key size without kernels */
  HEU_EKEY = 7          /* pk and sk are not paired / no secret key */
};

/* Message of the last failing call on this thread. */
const char *heu_last_error(void);

/* ---- context ---------------------------------------------------------- */
/* Creates a context on HIP device `device` with its own stream. */
int heu_ctx_create(int device, heu_ctx **out);
int heu_ctx_destroy(heu_ctx *ctx);
/* Makes subsequent work run on `stream` (a hipStream_t, used as given: NULL is
 * the device's null stream), e.g. the caller's framework stream.  A new
 * context uses a stream of its own. */
int heu_ctx_set_stream(heu_ctx *ctx, void *stream);
void *heu_ctx_stream(heu_ctx *ctx);
/* Waits for all enqueued work; returns HEU_ERANGE / HEU_ENOTINV if a kernel
 * flagged bad input since the last sync, then clears the flags. */
int heu_ctx_sync(heu_ctx *ctx);

/* Per-kernel timing (tracing): with enable != 0, every kernel launch of this
 * context is bracketed by HIP events on its stream; enabling or disabling
 * clears the records.  heu_ctx_profile_read sums the elapsed time of the
 * launches of kernel `name` ("encrypt_i64", "dec_half", "pow_i64", ...). */
int heu_ctx_profile(heu_ctx *ctx, int enable);
int heu_ctx_profile_read(heu_ctx *ctx, const char *name, double *total_ms, uint64_t *launches);

/* ---- key generation (paillier_zahlen/key_generator.cc:26-60) ------------ */
/* Writes n (key_bits/32 limbs), h_s (key_bits/16), p and q (key_bits/64 each).
 * seed == NULL draws from the OS RNG like the reference; otherwise every
 * random word comes from a counter stream keyed by *seed, so a seed always
 * gives the same key. */
int heu_keygen(uint32_t key_bits, const uint64_t *seed, uint32_t *n, uint32_t *h_s, uint32_t *p, uint32_t *q);

/* ---- keys (paillier_zahlen/public_key.cc:28-40, secret_key.cc:19-62) ---- */
/* n: n_words limbs; h_s: 2*n_words limbs.  p, q (prime_words limbs each) may
 * be NULL for a public-only key.  r_bits: largest bit length of the
 * randomness r the fixed-base table must cover (the reference draws
 * key_bits/2 bits, encryptor.cc:27); window_bits: table window (the
 * reference's kExpUnitBits = 10, public_key.cc:21), or 0 for the widest window
 * (<= 16) whose table fits HEU_TABLE_BYTES (default 4 GiB) per key. */
int heu_key_create(heu_ctx *ctx, const uint32_t *n, const uint32_t *h_s, uint32_t n_words, const uint32_t *p,
                   const uint32_t *q, uint32_t prime_words, uint32_t r_bits, uint32_t window_bits, heu_key **out);
int heu_key_destroy(heu_key *key);
int heu_key_info(const heu_key *key, uint32_t *key_bits, uint32_t *n_words, uint32_t *ct_words, uint32_t *has_sk);
/* Device bytes held by the key (constants + table). */
size_t heu_key_device_bytes(const heu_key *key);

/* Tunable per-key parameters (no counterpart in the reference, whose window
 * is the constant kExpUnitBits = 10, public_key.cc:21):
 *   HEU_PARAM_TABLE_WINDOW  fixed-base table window in bits (1..16); setting it
 *                           rebuilds the table (waits for the whole device);
 *   HEU_PARAM_DEC_LANES     lanes per bignum of the CRT decryption: 2 or 4 for
 *                           2048-bit keys, 4 or 8 for 3072-bit keys.
 * No work using the key may be in flight while a parameter is set. */
enum { HEU_PARAM_TABLE_WINDOW = 1, HEU_PARAM_DEC_LANES = 2 };
int heu_key_set_param(heu_key *key, int param, uint32_t value);
int heu_key_get_param(const heu_key *key, int param, uint32_t *value);

/* Autotuner: times the candidates of each parameter on this device with this
 * key (synthetic inputs) and keeps the fastest -- the table window (encryption
 * throughput), the decryption lanes per bignum (decryption throughput) and
 * the persistent work-list grid multiplier (occupancy) of the key's shapes,
 * which is set on the key's context.  A second decryption shape or a
 * multiplier above 1 is kept only when it is at least 2% faster.  `what` is a mask of HEU_TUNE_*; report
 * (may be NULL) receives a JSON object with every timing and the choices. */
enum { HEU_TUNE_TABLE = 1, HEU_TUNE_DEC_LANES = 2, HEU_TUNE_GRID = 4, HEU_TUNE_ALL = 7 };
int heu_autotune(heu_key *key, uint32_t what, char *report, size_t report_len);

/* ---- batch operations: device pointers, asynchronous ------------------- */
/* Encryptor::Encrypt (paillier_zahlen/encryptor.cc:35-49) with injected r:
 * ct[i] = (1 + n*m[i]) * h_s^r[i] mod n^2.  r: count x r_words limbs. */
int heu_encrypt_i64(heu_key *key, const int64_t *m, const uint32_t *r, uint32_t r_words, size_t count,
                    uint32_t *ct);
/* Same for plaintexts given as two's complement over m_words limbs; |m| <= n/2
 * (checked on the host copy `m_host`, which must hold the same values). */
int heu_encrypt(heu_key *key, const uint32_t *m, const uint32_t *m_host, uint32_t m_words, const uint32_t *r,
                uint32_t r_words, size_t count, uint32_t *ct);
/* Evaluator::Randomize (evaluator.cc:25-28): out = ct * h_s^r.  ct == NULL
 * gives Encryptor::EncryptZero (encryptor.cc:33). */
int heu_randomize(heu_key *key, const uint32_t *ct, const uint32_t *r, uint32_t r_words, size_t count,
                  uint32_t *out);
/* Decryptor::Decrypt (decryptor.cc:32-53), CRT.  out: count x out_words
 * limbs, the plaintext sign-extended or truncated to out_words; overflow[i]
 * (may be NULL) is set when it did not fit. */
int heu_decrypt(heu_key *key, const uint32_t *ct, size_t count, uint32_t *out, uint32_t out_words, uint8_t *overflow);
/* Evaluator::Add(ct, ct) (evaluator.cc:30-37): out = a * b mod n^2. */
int heu_add(heu_key *key, const uint32_t *a, const uint32_t *b, size_t count, uint32_t *out);
/* Evaluator::Add(ct, pt) (evaluator.cc:46-64): out = ct * (1 + n*pt). */
int heu_add_i64(heu_key *key, const uint32_t *ct, const int64_t *pt, size_t count, uint32_t *out);
/* Evaluator::Sub(ct, ct) (evaluator.cc:66-68): out = a * b^-1 mod n^2. */
int heu_sub(heu_key *key, const uint32_t *a, const uint32_t *b, size_t count, uint32_t *out);
/* Evaluator::Negate (evaluator.cc:82-91): out = a^-1 mod n^2. */
int heu_negate(heu_key *key, const uint32_t *a, size_t count, uint32_t *out);
/* Evaluator::Mul(ct, p) (evaluator.cc:99-123): out = ct^k mod n^2, k = 0 ->
 * 1, k = 1 -> ct, k < 0 -> inverse power. */
int heu_mul_i64(heu_key *key, const uint32_t *ct, const int64_t *k, size_t count, uint32_t *out);
/* Same for scalars of any length: |k| as k_words little-endian limbs per
 * element (k_mag) and its sign (k_neg[i] != 0 for k < 0; k_neg may be NULL). */
int heu_mul(heu_key *key, const uint32_t *ct, const uint32_t *k_mag, const uint8_t *k_neg, uint32_t k_words,
            size_t count, uint32_t *out);
/* numpy Evaluator::Sum per segment (numpy/evaluator.cc:355-373): out[s] =
 * prod ct[offsets[s] .. offsets[s+1]) mod n^2 (1 for an empty segment).
 * offsets: HOST array of segments+1 entries. */
int heu_sum(heu_key *key, const uint32_t *ct, const int64_t *offsets, size_t segments, uint32_t *out);
/* MatMul / dot per segment (numpy/evaluator.cc:209-292): out[s] =
 * prod ct[i]^w[i] over the segment.  w: device array, one weight per ct. */
int heu_dot_i64(heu_key *key, const uint32_t *ct, const int64_t *w, const int64_t *offsets, size_t segments,
                uint32_t *out);
/* Gathered forms for selections, bucket sums and matrix products
 * (numpy/evaluator.h:57-70 SelectSum, numpy/evaluator.cc:378-466
 * FeatureWiseBucketSum, :209-292 MatMul; pylib/numpy_binding/
 * extension_functions.cc:29-87 SelectSum / BatchSelectSum): positions j of
 * segment s, j in [offsets[s], offsets[s+1]) (HOST offsets), address
 * ciphertext idx[j] (device int64 indices into ct):
 *   heu_sum_gather: out[s] = prod ct[idx[j]]            (1 for an empty segment)
 *   heu_dot_gather: out[s] = prod ct[idx[j]]^w[j]       (w: device int64, per position)
 * The engine does not know how many ciphertexts ct holds.  So the caller
 * guarantees that every idx[j] addressed lies within it, as it does for the
 * offsets of heu_sum / heu_dot_i64 / heu_prefix_mul.  heu_amd/engine.py
 * checks both on the host before it calls. */
int heu_sum_gather(heu_key *key, const uint32_t *ct, const int64_t *idx, const int64_t *offsets, size_t segments,
                   uint32_t *out);
int heu_dot_gather(heu_key *key, const uint32_t *ct, const int64_t *idx, const int64_t *w, const int64_t *offsets,
                   size_t segments, uint32_t *out);
/* heu_sum for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments, kernels and results as
 * heu_sum; its scratch (chunk lists, the partial products of every reduction
 * level) lives in a fixed array of HEU_NOALLOC_SCRATCH_BYTES in each GPU's
 * memory, which that GPU holds from the first context created on it until the
 * process ends, so the call makes no device allocation.  A call that needs
 * more scratch fails with HEU_ERANGE before any launch (a 4096 x 256 matrix
 * summed along axis 0 needs about 40 MiB with 3072-bit keys).  Every call on
 * a GPU shares its array and the array keeps the last call's contents: calls
 * on one context run in stream order, but calls that overlap on one GPU (other
 * contexts or streams) race. */
#define HEU_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int heu_sum_noalloc(heu_key *key, const uint32_t *ct, const int64_t *offsets, size_t segments, uint32_t *out);
/* Inclusive prefix products within segments (the cumsum of
 * numpy/evaluator.cc:450-456): out[j] = ct[offsets[s]] * ... * ct[j] mod n^2
 * for j in segment s; out may alias ct. */
int heu_prefix_mul(heu_key *key, const uint32_t *ct, const int64_t *offsets, size_t segments, uint32_t *out);

/* Montgomery-form export / import for R = 2^(word_bits*ceil(bits(n^2)/word_bits))
 * (MontgomerySpace::MapIntoMSpace / MapBackToZSpace, encryptor.cc:46,
 * decryptor.cc:36). */
int heu_to_mont(heu_key *key, const uint32_t *ct, size_t count, uint32_t word_bits, uint32_t *out);
int heu_from_mont(heu_key *key, const uint32_t *ct, size_t count, uint32_t word_bits, uint32_t *out);

/* ---- engine form: ciphertexts resident in Montgomery form --------------- */
/* Callers that chain operations on the device may keep ciphertexts in the
 * engine's own form between calls: c R mod n^2 as 29-bit limbs in the
 * interleaved lane layout, *words (heu_key_mform_words) uint32 words per
 * element, lazily reduced (below 2 n^2, so bytes of equal values may differ;
 * heu_from_mform gives the canonical Z-space words).  In that form ct (+) ct
 * is one Montgomery product per element and a segment sum one per term,
 * where Z-space I/O costs a second product (or word conversions) per
 * element.  No counterpart in the reference, whose paillier_z::Ciphertext
 * also stays in a Montgomery space between operations (c_, encryptor.cc:46,
 * evaluator.cc:30-37) but with yacl's R. */
int heu_key_mform_words(const heu_key *key, uint32_t *words);
int heu_to_mform(heu_key *key, const uint32_t *ct, size_t count, uint32_t *out);
int heu_from_mform(heu_key *key, const uint32_t *mf, size_t count, uint32_t *out);
/* Evaluator::Add(ct, ct) (evaluator.cc:30-37) in engine form. */
int heu_add_mform(heu_key *key, const uint32_t *a, const uint32_t *b, size_t count, uint32_t *out);
/* heu_sum in engine form: inputs and the per-segment outputs (1 in engine
 * form for an empty segment) are engine-form values; HOST offsets. */
int heu_sum_mform(heu_key *key, const uint32_t *mf, const int64_t *offsets, size_t segments, uint32_t *out);

/* ---- wire format (host buffers; no device work) ------------------------ */
/* HEU's serialized form of ZPaillier objects, so engine ciphertexts and keys
 * interoperate with HEU peers.  Layouts follow the reference's writers:
 * paillier_z::Ciphertext is MSGPACK_DEFINE(c_), a msgpack array holding c_
 * (paillier_zahlen/ciphertext.h:37); PublicKey packs [n, h_s]
 * (public_key.h:64-76) and SecretKey [lambda, mu, p, q] (secret_key.h:64-80);
 * phe's variant wrapper appends the schema index as an 8-byte little-endian
 * size_t (phe/base/serializable_types.cc:50-70), HEU_SCHEMA_ZPAILLIER
 * (phe/base/schema.h:49).  c_ is the Montgomery form that heu_to_mont writes
 * (word_bits 64 for yacl's OpenSSL and GMP backends).  A BigInt is a msgpack
 * str of one sign byte (0 or 1) and the little-endian magnitude without high
 * zero bytes: yacl's own BigInt packing is not part of the reference, so that
 * inner byte layout is this engine's.  Writers return HEU_OK with the length
 * in *len, or HEU_ERANGE with the length needed when cap is too small (out
 * may then be NULL).  Readers check the layout, the schema index and that
 * values fit the given limbs (HEU_EINVAL / HEU_ERANGE otherwise). */
#define HEU_SCHEMA_ZPAILLIER 4
/* phe::Ciphertext::Serialize: `mont` holds words limbs of c R mod n^2. */
int heu_wire_put_ciphertext(const uint32_t *mont, uint32_t words, uint8_t *out, size_t cap, size_t *len);
int heu_wire_get_ciphertext(const uint8_t *in, size_t len, uint32_t *mont, uint32_t words);
/* numpy DenseMatrix<Ciphertext>::Serialize (numpy/matrix.h:221-265):
 * [rows, cols, ndim, [items]], every item the bytes of heu_wire_put_ciphertext
 * as a msgpack str.  mont: rows * cols values in column-major order (Eigen's
 * storage order).  The reader gives the shape (rows, cols, ndim may be NULL)
 * and, with mont != NULL, the values (cap_values >= rows * cols). */
int heu_wire_put_matrix(const uint32_t *mont, int64_t rows, int64_t cols, int64_t ndim, uint32_t words,
                        uint8_t *out, size_t cap, size_t *len);
int heu_wire_get_matrix(const uint8_t *in, size_t len, int64_t *rows, int64_t *cols, int64_t *ndim, uint32_t *mont,
                        uint32_t words, size_t cap_values);
/* DenseMatrix<Ciphertext>::Serialize4Ic / LoadFromIc (numpy/matrix.cc:39-111):
 * the interconnection form, a protobuf DataExchangeProtocol with scalar_type
 * OBJECT, scalar_type_name "ciphertext" and an NdArray of the shape (no entry
 * for ndim 0, [rows] for 1, [rows, cols] for 2) and the items, each the bytes
 * of heu_wire_put_ciphertext, column-major as above.  The reader ignores the
 * type name and unknown fields and requires scalar_type OBJECT, an NdArray and
 * as many items as the shape holds.  The .proto belongs to org_interconnection
 * 0.0.1, which the reference does not vendor: the field numbers and the enum
 * value are listed in wire.cpp. */
int heu_wire_put_matrix_ic(const uint32_t *mont, int64_t rows, int64_t cols, int64_t ndim, uint32_t words,
                           uint8_t *out, size_t cap, size_t *len);
int heu_wire_get_matrix_ic(const uint8_t *in, size_t len, int64_t *rows, int64_t *cols, int64_t *ndim, uint32_t *mont,
                           uint32_t words, size_t cap_values);
/* phe::PublicKey / phe::SecretKey for ZPaillier.  The secret key's lambda =
 * (p-1)(q-1)/2 and mu = lambda^-1 mod n (key_generator.cc:26-60) are computed
 * from p and q when written and checked against them when read.  Readers
 * report the word count (*n_words = words of n; *prime_words = words of the
 * larger prime) and write the values when the buffers are given (h_s gets
 * 2 * cap_words limbs). */
int heu_wire_put_public_key(const uint32_t *n, uint32_t n_words, const uint32_t *h_s, uint8_t *out, size_t cap,
                            size_t *len);
int heu_wire_get_public_key(const uint8_t *in, size_t len, uint32_t *n, uint32_t *h_s, uint32_t cap_words,
                            uint32_t *n_words);
int heu_wire_put_secret_key(const uint32_t *p, const uint32_t *q, uint32_t prime_words, uint8_t *out, size_t cap,
                            size_t *len);
int heu_wire_get_secret_key(const uint8_t *in, size_t len, uint32_t *p, uint32_t *q, uint32_t cap_words,
                            uint32_t *prime_words);

/* ---- multi-task batches ------------------------------------------------ */
/* A batch runs many tasks -- vectors of any length, on any mix of resident
 * 2048- and 3072-bit keys -- as one dispatch.  The elementwise work of every
 * task runs in persistent work-list launches, one per (key size, modulus
 * shape): workgroups pull units of elements from a shared counter, so short
 * and long vectors balance across the chip.  Dependent follow-up work (CRT
 * recombination, batch inversion for negative scalars, upper reduction levels)
 * follows in task order.  Tasks between barriers (HEU_OP_BARRIER) are
 * independent: none may read another's output.  With HEU_BATCH_GRAPH the whole sequence is captured
 * once into a hipGraph and every run replays it (one hipGraphLaunch).
 *
 * This is the shape HEU's numpy layer hands a vectorized backend -- one span
 * per yacl worker chunk, numpy/encryptor.cc:42-52, numpy/evaluator.cc:120-186
 * -- gathered across operations and keys; it has no single counterpart in the
 * reference, whose GPU backend makes one synchronous call per op
 * (paillier_gpu/gpulib/gpupaillier.h:17-91). */
typedef struct heu_batch heu_batch;

enum {
  HEU_OP_ENCRYPT_I64 = 1, /* a: int64 m[count], b: r (aux limbs each); out: ct */
  HEU_OP_RANDOMIZE = 2,   /* a: ct or NULL (EncryptZero), b: r (aux limbs each); out: ct */
  HEU_OP_ADD = 3,         /* a, b: ct; out: a*b */
  HEU_OP_ADD_I64 = 4,     /* a: ct, b: int64 pt; out: ct*(1+n*pt) */
  HEU_OP_MUL_I64 = 5,     /* a: ct, b: int64 scalar; out: ct^k (inverse power for k < 0) */
  HEU_OP_DECRYPT = 6,     /* a: ct; out: plaintext, aux limbs two's complement each */
  HEU_OP_SUM = 7,         /* a: ct; offsets/segments; out: one ct per segment */
  HEU_OP_DOT_I64 = 8,     /* a: ct, b: int64 weights; offsets/segments; out: one ct per segment */
  HEU_OP_BARRIER = 9      /* no operands: tasks after it see the outputs of every task before it */
};

typedef struct heu_task {
  heu_key *key;
  int32_t op;              /* HEU_OP_* */
  uint32_t aux;            /* r limbs (ENCRYPT_I64, RANDOMIZE) or output limbs (DECRYPT) */
  const void *a;           /* device inputs as listed per op */
  const void *b;
  size_t count;            /* elements (for SUM / DOT: offsets[segments] - offsets[0] is ignored) */
  const int64_t *offsets;  /* SUM / DOT: HOST array of segments + 1 element offsets into a */
  size_t segments;
  void *out;               /* device output */
} heu_task;

#define HEU_BATCH_GRAPH 1u /* capture into a hipGraph at the first run, replay afterwards */

/* Plans a batch: sizes and allocates its own device workspace, uploads its
 * work lists, and copies `tasks` (offsets included).  Device pointers in the
 * tasks are used as given at every run, and every key must outlive the batch. */
int heu_batch_create(heu_ctx *ctx, const heu_task *tasks, size_t ntasks, uint32_t flags, heu_batch **out);
/* Enqueues one run on the context's stream (asynchronous; errors found by
 * kernels surface at heu_ctx_sync). */
int heu_batch_run(heu_batch *batch);
int heu_batch_destroy(heu_batch *batch);
/* Kernel launches per run (graph nodes when captured), work units in the
 * persistent launches, and device bytes held by the batch. */
int heu_batch_stats(const heu_batch *batch, uint64_t *launches, uint64_t *units, size_t *device_bytes);

/* Persistent launch width: workgroups = multiplier x resident workgroups per
 * CU x CUs (at most one per work unit), multiplier in [1, 64], default 1.
 * Applies to batches planned afterwards; the autotuner sets it per shape.  The
 * request server budgets by what runs at once (multiplier 1) either way. */
int heu_ctx_set_grid_multiplier(heu_ctx *ctx, uint32_t key_bits, int decrypt_shape, uint32_t multiplier);

/* ---- request server (continuous batching) ------------------------------ */
/* A scheduler thread admits queued requests (one heu_task each, device
 * operands, any resident key) into batches while device workspace is free,
 * runs up to max_inflight batches at once on streams of its own, and retires
 * them as they finish: requests that arrive while others run are batched and
 * started as soon as a stream slot and workspace are free, so their work
 * overlaps on the device.  Elementwise requests are cut into parts of a
 * fraction of a round of the device (the work units its persistent launch runs
 * at once), and every batch takes parts round-robin across the active requests,
 * shortest remaining work first, so
 * a long request shares the device with the short ones that arrive during it
 * instead of blocking them (sums and dots stay whole).  A request that has
 * waited longer than urgent_ms goes first, which bounds how long a long
 * request can be held back by a stream of short ones.  A request completes
 * when its last part does, and its result is independent of what it was
 * batched with.  HEU has no counterpart; it serves the request-loop shape of
 * the path (keys and tables resident, requests of mixed sizes, keys and ops). */
typedef struct heu_server heu_server;

typedef struct heu_server_config {
  size_t workspace_bytes;   /* initial device workspace for in-flight batches (0: 8 GiB) */
  uint32_t max_inflight;    /* batches on the device at once (0: 2, at most 8) */
  uint32_t max_batch_tasks; /* request parts per batch (0: 512) */
  uint32_t batch_rounds;    /* work-list units per batch and kernel shape, in rounds of the
                               units the device runs at once (0: 1) */
  uint32_t parts_per_round; /* elementwise requests are cut into parts of 1/parts_per_round of a round (0: 8) */
  uint32_t urgent_ms;       /* a request waiting longer than this goes ahead of every newer one and takes
                               its shape's whole room, oldest first (0: 150; UINT32_MAX: never) */
  size_t max_workspace_bytes; /* a batch that does not fit the free workspace grows it by a new device
                               segment while hipMemGetInfo shows more than a reserve (5% of the device, at
                               least 2 GiB) free; the workspace stays within this (0: no other limit) */
} heu_server_config;

typedef struct heu_server_counters {
  uint64_t submitted, completed, failed, batches;
  uint64_t max_batch_requests; /* most requests seen in one batch */
  size_t peak_workspace_bytes;
  size_t workspace_bytes;       /* device workspace held now (initial plus grown segments) */
} heu_server_counters;

int heu_server_create(heu_ctx *ctx, const heu_server_config *cfg, heu_server **out);
/* Queues one request; *ticket identifies it.  Thread-safe.  The task is
 * checked here (bad arguments fail the call); keys and buffers must stay
 * valid until the request completes. */
int heu_server_submit(heu_server *srv, const heu_task *task, uint64_t *ticket);
/* Waits for a request (timeout_ms < 0: no limit).  HEU_OK: its output is
 * written and *latency_ms (may be NULL) is submit -> completion on the host
 * clock; HEU_EINVAL for an unknown ticket; HEU_EDEVICE on timeout or if the
 * request could not be run (e.g. it alone exceeds the workspace).  A waited
 * ticket is forgotten. */
int heu_server_wait(heu_server *srv, uint64_t ticket, double timeout_ms, double *latency_ms);
int heu_server_stats(heu_server *srv, heu_server_counters *stats);
/* Drains every queued and running request, then stops. */
int heu_server_destroy(heu_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* HEU_PAILLIER_H_ */
