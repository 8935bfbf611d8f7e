// engine.cpp -- host side of the C ABI declared in include/heu_paillier.h.
//
// Contexts own a stream, a device error word and a growable scratch pool.
// Keys hold every per-key constant on the device: the R29 moduli and
// Montgomery constants for n^2, p^2 and q^2, the CRT constants, the window
// schedules of p-1 and q-1, and the fixed-base table of h_s (built on the
// device).  Per-key constants are computed once on the host with OpenSSL
// BIGNUM at key admission; no per-element work runs on the host.
#include <hip/hip_runtime.h>
#include <openssl/bn.h>

#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/heu_paillier.h"
#include "engine_internal.hpp"
#include "kernels.hpp"
#include "mp_group.hpp"

using heu_hip::Chunk;
using heu_hip::CombArgs;
using heu_hip::DecArgs;
using heu_hip::ModArgs;
using heu_hip::Ops;
using heu_hip::PubArgs;

namespace heu_engine {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

static void ctx_free(heu_ctx* ctx) {
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->d_err) (void)hipFree(ctx->d_err);
  ctx->scratch.alloc(0);
  ctx->meta.alloc(0);
  ctx->prof_clear();
  for (hipEvent_t e : ctx->prof_pool) (void)hipEventDestroy(e);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  delete ctx;
}

void ctx_retain(heu_ctx* ctx) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  ++ctx->refs;
}

void ctx_release(heu_ctx* ctx) {
  bool free_now;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    free_now = --ctx->refs == 0 && ctx->closing;
  }
  if (free_now) ctx_free(ctx);
}

std::vector<long> inv_levels(long count, long chunk0) {
  std::vector<long> cnt{count};
  while (cnt.size() < 2 || cnt.back() > 1) {
    const long c = inv_chunk(cnt.size() - 1, chunk0);
    cnt.push_back((cnt.back() + c - 1) / c);
  }
  return cnt;
}

// Groups of the n^2 shape that fill the chip about twice over.
static long fill_groups(heu_ctx* ctx, const heu_hip::Ops* o) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus <= 0)
    cus = 1;
  return 2L * cus * (heu_hip::kBlock / o->n_L);
}

long reduce_chunk0(heu_ctx* ctx, const heu_hip::Ops* o, bool dot, long total) {
  const long target = fill_groups(ctx, o);
  const long cap = dot ? kReduceChunkDot0 : kReduceChunk0;
  return std::max(1L, std::min(cap, (total + target - 1) / target));
}

long inv_chunk0(heu_ctx* ctx, const heu_hip::Ops* o, long count) {
  const long target = fill_groups(ctx, o);
  return std::max(kInvChunk0Min, std::min(kInvChunk0, (count + target - 1) / target));
}

std::vector<heu_hip::Chunk> make_chunks(const std::vector<long>& seg_start, const std::vector<long>& seg_len, long chunk,
                               std::vector<long>* seg_first_chunk, std::vector<long>* seg_nchunks) {
  std::vector<heu_hip::Chunk> ch;
  size_t segs = seg_len.size();
  seg_first_chunk->assign(segs, 0);
  seg_nchunks->assign(segs, 0);
  for (size_t s = 0; s < segs; ++s) {
    (*seg_first_chunk)[s] = (long)ch.size();
    long len = seg_len[s];
    if (len == 0) {
      ch.push_back(heu_hip::Chunk{seg_start[s], 0});
    } else {
      for (long off = 0; off < len; off += chunk) ch.push_back(heu_hip::Chunk{seg_start[s] + off, std::min(chunk, len - off)});
    }
    (*seg_nchunks)[s] = (long)ch.size() - (*seg_first_chunk)[s];
  }
  return ch;
}

}  // namespace heu_engine

using namespace heu_engine;

namespace {

// ---------------------------------------------------------------------------
// BIGNUM helpers
// ---------------------------------------------------------------------------
struct BnCtx {
  BN_CTX* c = BN_CTX_new();
  ~BnCtx() { BN_CTX_free(c); }
};
struct Bn {
  BIGNUM* b = BN_new();
  Bn() = default;
  Bn(const Bn&) = delete;
  ~Bn() { BN_free(b); }
  operator BIGNUM*() const { return b; }
};

void bn_from_words(BIGNUM* r, const uint32_t* w, size_t words) {
  BN_lebin2bn(reinterpret_cast<const unsigned char*>(w), (int)(words * 4), r);
}
std::vector<uint32_t> bn_to_words(const BIGNUM* a, size_t words) {
  std::vector<uint32_t> out(words, 0);
  BN_bn2lebinpad(a, reinterpret_cast<unsigned char*>(out.data()), (int)(words * 4));
  return out;
}
uint32_t bn_limb(const BIGNUM* a, int j) {  // 29-bit limb j
  uint32_t v = 0;
  for (int b = 0; b < heu_hip::kLimbBits; ++b)
    if (BN_is_bit_set(a, j * heu_hip::kLimbBits + b)) v |= 1u << b;
  return v;
}
// R29 interleaved layout for a group of L lanes x K limbs.
std::vector<uint32_t> r29_interleaved(const BIGNUM* a, int L, int K) {
  std::vector<uint32_t> out((size_t)L * K);
  for (int l = 0; l < L; ++l)
    for (int t = 0; t < K; ++t) out[(size_t)t * L + l] = bn_limb(a, l * K + t);
  return out;
}
std::vector<uint32_t> r29_plain(const BIGNUM* a, int NL) {
  std::vector<uint32_t> out(NL);
  for (int j = 0; j < NL; ++j) out[j] = bn_limb(a, j);
  return out;
}
// -m^-1 mod 2^bits for odd m
uint64_t neg_inv_pow2(const BIGNUM* m, int bits) {
  uint64_t m0 = BN_get_word(m) ;  // low bits of m are enough
  if (BN_num_bits(m) > 64) {
    Bn t;
    BN_copy(t, m);
    BN_mask_bits(t, 64);
    m0 = BN_get_word(t);
  }
  uint64_t inv = 1;  // Newton: inv = inv * (2 - m0 * inv)
  for (int i = 0; i < 7; ++i) inv *= 2 - m0 * inv;
  uint64_t mask = bits == 64 ? ~0ull : ((1ull << bits) - 1);
  return (0 - inv) & mask;
}

// Sliding-window schedule (left to right, odd powers, window `w`) of e:
// op = (table index + 1) << 16 | squarings before the multiply.
std::vector<uint32_t> window_schedule(const BIGNUM* e, int w) {
  std::vector<uint32_t> ops;
  int i = BN_num_bits(e) - 1;
  auto window_at = [&](int top, int* low) {
    int j = std::max(top - w + 1, 0);
    while (!BN_is_bit_set(e, j)) ++j;
    uint32_t v = 0;
    for (int b = top; b >= j; --b) v = (v << 1) | (uint32_t)BN_is_bit_set(e, b);
    *low = j;
    return v;
  };
  int j;
  uint32_t v = window_at(i, &j);
  ops.push_back((((v - 1) / 2) + 1) << 16);
  i = j - 1;
  uint32_t pending = 0;
  while (i >= 0) {
    if (!BN_is_bit_set(e, i)) {
      ++pending;
      --i;
      continue;
    }
    v = window_at(i, &j);
    pending += (uint32_t)(i - j + 1);
    ops.push_back(((((v - 1) / 2) + 1) << 16) | pending);
    pending = 0;
    i = j - 1;
  }
  if (pending) ops.push_back(pending);
  return ops;
}

}  // namespace

namespace {

int upload(DevBuf& buf, const std::vector<uint32_t>& v, hipStream_t st, size_t* total) {
  int rc = buf.alloc(v.size() * 4);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(buf.p, v.data(), v.size() * 4, hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  *total += v.size() * 4;
  return HEU_OK;
}

// Montgomery constants of modulus m for a group of L lanes x K limbs.
int upload_mod(const BIGNUM* m, int L, int K, DevBuf& dm, DevBuf& done, DevBuf& dr2, DevBuf* dr3, uint32_t* mprime,
               hipStream_t st, size_t* total, BN_CTX* ctx) {
  const int NL = L * K;
  Bn R, r;
  BN_set_bit(R, heu_hip::kLimbBits * NL);
  int rc = upload(dm, r29_interleaved(m, L, K), st, total);
  if (rc) return rc;
  BN_nnmod(r, R, m, ctx);
  if ((rc = upload(done, r29_interleaved(r, L, K), st, total))) return rc;
  BN_mod_mul(r, r, R, m, ctx);  // R^2 mod m
  if ((rc = upload(dr2, r29_interleaved(r, L, K), st, total))) return rc;
  if (dr3) {
    BN_mod_mul(r, r, R, m, ctx);
    if ((rc = upload(*dr3, r29_interleaved(r, L, K), st, total))) return rc;
  }
  *mprime = (uint32_t)neg_inv_pow2(m, heu_hip::kLimbBits);
  return HEU_OK;
}

// Builds (or rebuilds) the fixed-base table of h_s for a window of
// `window_bits` (public_key.cc:36-39 MakeBaseTable): T[w][d] = h_s^(d 2^(unit w)) R
// for every window w covering key->r_bits, built on the device by levels.
int build_table(heu_key* key, uint32_t window_bits) {
  heu_ctx* ctx = key->ctx;
  const Ops* ops = key->ops;
  hipStream_t st = ctx->stream;
  BnCtx bc;
  Bn b, n2;
  bn_from_words(b, key->hs_host.data(), key->hs_host.size());
  bn_from_words(n2, key->n2_host.data(), key->n2_host.size());
  const int windows = (int)((key->r_bits + window_bits - 1) / window_bits);
  std::vector<uint32_t> bases((size_t)windows * key->ct_words);
  for (int w = 0; w < windows; ++w) {
    auto v = bn_to_words(b, key->ct_words);
    std::copy(v.begin(), v.end(), bases.begin() + (size_t)w * key->ct_words);
    for (uint32_t s = 0; s < window_bits; ++s) BN_mod_sqr(b, b, n2, bc.c);
  }
  DevBuf dbases;
  size_t dummy = 0;
  int rc = upload(dbases, bases, st, &dummy);
  if (rc) return rc;
  HIP_TRY(hipDeviceSynchronize());  // no launch may still read the old table
  key->device_bytes -= key->table.bytes;
  const size_t tbytes = ((size_t)windows << window_bits) * key->nl_n * 4;
  if ((rc = key->table.alloc(tbytes))) return rc;
  key->device_bytes += tbytes;
  key->unit = (int)window_bits;
  key->windows = windows;
  for (int level = 0; level <= (int)window_bits; ++level)
    LAUNCH(ctx, "table_level", ops->table_level(st, key->mod_n2(), dbases.as<uint32_t>(), key->table.as<uint32_t>(),
                                                windows, key->unit, level));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  return HEU_OK;
}

// Carves 256-aligned parts from the context's scratch pool, starting at
// `base`.  The pool is reallocated (after draining the stream) only by a
// top-level carve (base == 0); nested carves must fit what the caller reserved.
// Over a fixed buffer (`fixed`, `fixed_bytes`) nothing is ever allocated: a
// carve that does not fit fails.
struct Scratch {
  heu_ctx* ctx;
  size_t base, total = 0;
  char* fixed = nullptr;
  size_t fixed_bytes = 0;
  explicit Scratch(heu_ctx* c, size_t b = 0) : ctx(c), base(b) {}
  Scratch(heu_ctx* c, char* buf, size_t bytes) : ctx(c), base(0), fixed(buf), fixed_bytes(bytes) {}
  size_t add(size_t bytes) {
    size_t off = base + total;
    total += (bytes + 255) & ~(size_t)255;
    return off;
  }
  size_t end() const { return base + total; }
  int commit() {
    if (fixed) {
      if (end() > fixed_bytes)
        return fail(HEU_ERANGE, "the call needs %zu bytes of scratch; the fixed buffer holds %zu", end(),
                    fixed_bytes);
      return HEU_OK;
    }
    if (end() > ctx->scratch.bytes) {
      if (base != 0) return fail(HEU_EINVAL, "internal: nested scratch exceeds the reservation");
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      int rc = ctx->scratch.alloc(std::max(end(), ctx->scratch.bytes * 2));
      if (rc) return rc;
    }
    return HEU_OK;
  }
  template <class T>
  T* at(size_t off) const {
    return reinterpret_cast<T*>((fixed ? fixed : static_cast<char*>(ctx->scratch.p)) + off);
  }
};

// Elements per slice of ops whose scratch grows with the batch (window tables).
constexpr long kSliceElems = 1l << 17;

int check_key(heu_key* key) {
  if (!key || !key->ops) return fail(HEU_EINVAL, "null key");
  return HEU_OK;
}

// ---------------------------------------------------------------------------
// Batch inversion of selected Z-space elements, in place (Montgomery's trick
// over chunks, levels up to a single value inverted by extended Euclid).
// ---------------------------------------------------------------------------
// Scratch bytes batch_invert carves for `count` elements.
size_t inv_scratch_bytes(const heu_key* key, long count) {
  const size_t NLB = (size_t)key->nl_n * 4;
  auto cnt = inv_levels(count, inv_chunk0(key->ctx, key->ops, count));
  size_t b = (((size_t)count * 4 + 255) & ~(size_t)255) + 256;  // index list and its count
  for (size_t l = 0; l < cnt.size(); ++l) {
    size_t part = (cnt[l] * NLB + 255) & ~(size_t)255;
    b += (l + 1 < cnt.size() ? 3 : 2) * part;
  }
  return b;
}

int batch_invert(heu_key* key, const uint32_t* in, const uint8_t* sel, long count, uint32_t* out, size_t base) {
  heu_ctx* ctx = key->ctx;
  const Ops* o = key->ops;
  const ModArgs mod = key->mod_n2();
  const size_t NLB = (size_t)key->nl_n * 4;
  if (count <= 0) return HEU_OK;
  const long chunk0 = inv_chunk0(ctx, o, count);
  std::vector<long> cnt = inv_levels(count, chunk0);
  const int levels = (int)cnt.size() - 1;  // cnt[levels] == 1
  Scratch sc(ctx, base);
  // selected elements (always in place here) run from a packed index list
  const bool compact = sel != nullptr && in == out;
  const size_t idx_off = sc.add((size_t)count * 4), nsel_off = sc.add(4);
  std::vector<size_t> pre(levels + 1), tot(levels + 1), inv(levels + 1);
  for (int l = 0; l <= levels; ++l) {
    if (l < levels) pre[l] = sc.add(cnt[l] * NLB);
    tot[l] = sc.add(cnt[l] * NLB);  // values of level l (level 0 unused)
    inv[l] = sc.add(cnt[l] * NLB);  // inverses of level l values (level 0 unused)
  }
  int rc = sc.commit();
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  heu_hip::InvSel s0{sel, nullptr, nullptr};
  if (compact) {
    s0 = heu_hip::InvSel{nullptr, sc.at<uint32_t>(idx_off), sc.at<uint32_t>(nsel_off)};
    LAUNCH(ctx, "inv_compact", o->inv_compact(st, sel, count, sc.at<uint32_t>(idx_off), sc.at<uint32_t>(nsel_off)));
  }
  const heu_hip::InvSel none{nullptr, nullptr, nullptr};
  // up: level l reduces cnt[l] values into cnt[l+1] totals
  for (int l = 0; l < levels; ++l) {
    const uint32_t* src = l == 0 ? in : sc.at<uint32_t>(tot[l]);
    LAUNCH(ctx, "inv_prefix", o->inv_prefix(st, mod, l == 0, src, l == 0 ? s0 : none, cnt[l], inv_chunk(l, chunk0),
                                            sc.at<uint32_t>(pre[l]), sc.at<uint32_t>(tot[l + 1])));
  }
  // the single top value
  HIP_TRY(hipMemcpyAsync(sc.at<uint32_t>(inv[levels]), sc.at<uint32_t>(tot[levels]), NLB, hipMemcpyDeviceToDevice, st));
  LAUNCH(ctx, "inv_single", o->inv_single(st, mod, key->n2_words.as<uint32_t>(), sc.at<uint32_t>(inv[levels]), ctx->d_err));
  // down
  for (int l = levels - 1; l >= 0; --l) {
    const uint32_t* src = l == 0 ? in : sc.at<uint32_t>(tot[l]);
    uint32_t* dst = l == 0 ? out : sc.at<uint32_t>(inv[l]);
    LAUNCH(ctx, "inv_down", o->inv_down(st, mod, l == 0, src, l == 0 ? s0 : none, cnt[l], inv_chunk(l, chunk0),
                                        sc.at<uint32_t>(pre[l]), sc.at<uint32_t>(inv[l + 1]), dst));
  }
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

int upload_meta(heu_ctx* ctx, Scratch& sc, size_t off, const void* host, size_t bytes) {
  HIP_TRY(hipMemcpyAsync(sc.at<char>(off), host, bytes, hipMemcpyHostToDevice, ctx->stream));
  // the host vector may die after return: make the copy complete now
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return HEU_OK;
}

// Generic segmented reduction.  dot: weights given (device), else plain product.
// With `fixed` set, a plain product carves its scratch from that buffer only.
// With `mform`, a plain product reads and writes engine-form values.
int segmented(heu_key* key, const uint32_t* ct, const int64_t* w, const int64_t* offsets, size_t segments,
              uint32_t* out, const int64_t* idx = nullptr, char* fixed = nullptr, size_t fixed_bytes = 0,
              bool mform = false) {
  heu_ctx* ctx = key->ctx;
  const Ops* o = key->ops;
  const ModArgs mod = key->mod_n2();
  const size_t NLB = (size_t)key->nl_n * 4;
  const bool dot = w != nullptr;
  if (segments == 0) return HEU_OK;
  if ((fixed || mform) && dot) return fail(HEU_EINVAL, "internal: a dot product needs Z-space inputs and the pool");
  if (mform && idx) return fail(HEU_EINVAL, "internal: engine-form sums take no index list");
  std::vector<long> start(segments), len(segments);
  for (size_t s = 0; s < segments; ++s) {
    start[s] = offsets[s];
    len[s] = offsets[s + 1] - offsets[s];
    if (len[s] < 0) return fail(HEU_EINVAL, "offsets must be non-decreasing");
  }
  long total = 0;
  for (long l : len) total += l;
  const long chunk0 = reduce_chunk0(ctx, o, dot, total), chunkU = kReduceChunkUpper;
  hipStream_t st = ctx->stream;
  // Plan all levels first (host), then carve scratch once.
  struct Level {
    std::vector<Chunk> ch;
    std::vector<long> first, nch;
    size_t ch_off, pos_off, neg_off;
  };
  std::vector<Level> lv;
  {
    Level L0;
    L0.ch = make_chunks(start, len, chunk0, &L0.first, &L0.nch);
    lv.push_back(std::move(L0));
    while (true) {
      const Level& prev = lv.back();
      bool done = true;
      for (size_t s = 0; s < segments; ++s) done &= prev.nch[s] == 1;
      if (done) break;
      Level n;
      n.ch = make_chunks(prev.first, prev.nch, chunkU, &n.first, &n.nch);
      lv.push_back(std::move(n));
    }
  }
  Scratch sc = fixed ? Scratch(ctx, fixed, fixed_bytes) : Scratch(ctx);
  for (auto& l : lv) {
    l.ch_off = sc.add(l.ch.size() * sizeof(Chunk));
    l.pos_off = sc.add(l.ch.size() * NLB);
    l.neg_off = dot ? sc.add(l.ch.size() * NLB) : 0;
  }
  size_t tab_off = dot ? sc.add(lv[0].ch.size() * (NLB << heu_hip::kPowWindowBits)) : 0;
  size_t top_pos = sc.add(segments * NLB), top_neg = dot ? sc.add(segments * NLB) : 0;
  size_t nz_off = dot ? sc.add(segments * (size_t)key->ct_words * 4) : 0;
  size_t inv_base = dot ? sc.add(inv_scratch_bytes(key, (long)segments)) : 0;
  int rc = sc.commit();
  if (rc) return rc;
  for (auto& l : lv)
    if ((rc = upload_meta(ctx, sc, l.ch_off, l.ch.data(), l.ch.size() * sizeof(Chunk)))) return rc;
  // level 0
  if (dot)
    LAUNCH(ctx, "reduce_dot0", o->reduce_dot0(st, mod, ct, w, sc.at<Chunk>(lv[0].ch_off), (long)lv[0].ch.size(), sc.at<uint32_t>(lv[0].pos_off),
                   sc.at<uint32_t>(lv[0].neg_off), sc.at<uint32_t>(tab_off), idx));
  else
    LAUNCH(ctx, "reduce_sum", o->reduce_sum(st, mod, !mform, ct, sc.at<Chunk>(lv[0].ch_off), (long)lv[0].ch.size(),
                                            sc.at<uint32_t>(lv[0].pos_off), idx));
  for (size_t i = 1; i < lv.size(); ++i) {
    if (dot) {
      LAUNCH(ctx, "reduce_mont", o->reduce_mont(st, mod, sc.at<uint32_t>(lv[i - 1].pos_off), sc.at<Chunk>(lv[i].ch_off), (long)lv[i].ch.size(),
                     sc.at<uint32_t>(lv[i].pos_off)));
      LAUNCH(ctx, "reduce_mont", o->reduce_mont(st, mod, sc.at<uint32_t>(lv[i - 1].neg_off), sc.at<Chunk>(lv[i].ch_off), (long)lv[i].ch.size(),
                     sc.at<uint32_t>(lv[i].neg_off)));
    } else {
      LAUNCH(ctx, "reduce_sum", o->reduce_sum(st, mod, false, sc.at<uint32_t>(lv[i - 1].pos_off), sc.at<Chunk>(lv[i].ch_off),
                    (long)lv[i].ch.size(), sc.at<uint32_t>(lv[i].pos_off), nullptr));
    }
  }
  // the last level has one chunk per segment, in segment order
  const Level& top = lv.back();
  if (mform) {  // a product of engine-form values is in engine form: done
    HIP_TRY(hipMemcpyAsync(out, sc.at<char>(top.pos_off), segments * NLB, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipGetLastError());
    return HEU_OK;
  }
  HIP_TRY(hipMemcpyAsync(sc.at<char>(top_pos), sc.at<char>(top.pos_off), segments * NLB, hipMemcpyDeviceToDevice, st));
  if (!dot) {
    LAUNCH(ctx, "seg_finish", o->seg_finish(st, mod, sc.at<uint32_t>(top_pos), (long)segments, out));
  } else {
    HIP_TRY(
        hipMemcpyAsync(sc.at<char>(top_neg), sc.at<char>(top.neg_off), segments * NLB, hipMemcpyDeviceToDevice, st));
    // N -> Z-space, invert all N in one batch, result = P * N^-1
    uint32_t* nz = sc.at<uint32_t>(nz_off);
    LAUNCH(ctx, "seg_finish", o->seg_finish(st, mod, sc.at<uint32_t>(top_neg), (long)segments, nz));
    if ((rc = batch_invert(key, nz, nullptr, (long)segments, nz, inv_base))) return rc;
    LAUNCH(ctx, "mont_times_z", o->mont_times_z(st, mod, sc.at<uint32_t>(top_pos), nz, (long)segments, out));
  }
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* heu_last_error(void) { return g_last_error.c_str(); }

int heu_ctx_create(int device, heu_ctx** out) {
  if (!out) return fail(HEU_EINVAL, "null out");
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(HEU_EINVAL, "device %d not present (%d devices)", device, n);
  HIP_TRY(hipSetDevice(device));
  // the device's fixed scratch of heu_sum_noalloc; a failure is kept for that
  // entry point to report and does not stop the context
  (void)noalloc_setup(device);
  auto ctx = new heu_ctx();
  ctx->device = device;
  hipError_t e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
  ctx->stream = ctx->own_stream;
  if (e == hipSuccess) e = hipMalloc(&ctx->d_err, sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemset(ctx->d_err, 0, sizeof(uint32_t));
  if (e != hipSuccess) {
    ctx_free(ctx);
    return fail(e == hipErrorOutOfMemory ? HEU_ENOMEM : HEU_EDEVICE, "context set-up: %s", hipGetErrorString(e));
  }
  *out = ctx;
  return HEU_OK;
}

int heu_ctx_destroy(heu_ctx* ctx) {
  if (!ctx) return HEU_OK;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->closing = true;
    if (ctx->refs > 0) return HEU_OK;  // freed with the last key / batch / server
  }
  ctx_free(ctx);
  return HEU_OK;
}

int heu_ctx_profile(heu_ctx* ctx, int enable) {
  if (!ctx) return fail(HEU_EINVAL, "null ctx");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->prof_clear();
  ctx->prof = enable != 0;
  return HEU_OK;
}

int heu_ctx_profile_read(heu_ctx* ctx, const char* name, double* total_ms, uint64_t* launches) {
  if (!ctx || !name || !total_ms || !launches) return fail(HEU_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  double t = 0;
  uint64_t n = 0;
  for (auto& r : ctx->prof_recs) {
    if (strcmp(r.name, name) != 0) continue;
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, r.start, r.stop));
    t += ms;
    ++n;
  }
  *total_ms = t;
  *launches = n;
  return HEU_OK;
}

int heu_ctx_set_stream(heu_ctx* ctx, void* stream) {
  if (!ctx) return fail(HEU_EINVAL, "null ctx");
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->stream = static_cast<hipStream_t>(stream);
  return HEU_OK;
}

void* heu_ctx_stream(heu_ctx* ctx) { return ctx ? static_cast<void*>(ctx->stream) : nullptr; }

int heu_ctx_sync(heu_ctx* ctx) {
  if (!ctx) return fail(HEU_EINVAL, "null ctx");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  uint32_t flags = 0;
  HIP_TRY(hipMemcpy(&flags, ctx->d_err, sizeof(flags), hipMemcpyDeviceToHost));
  if (flags) HIP_TRY(hipMemset(ctx->d_err, 0, sizeof(uint32_t)));
  if (flags & heu_hip::kErrNotInvertible) return fail(HEU_ENOTINV, "a value has no inverse mod n^2");
  if (flags & heu_hip::kErrRange) return fail(HEU_ERANGE, "randomness r exceeds the key's table range");
  return HEU_OK;
}

int heu_key_create(heu_ctx* ctx, const uint32_t* n_w, const uint32_t* hs_w, uint32_t n_words, const uint32_t* p_w,
                   const uint32_t* q_w, uint32_t prime_words, uint32_t r_bits, uint32_t window_bits, heu_key** out) {
  if (!ctx || !n_w || !hs_w || !out) return fail(HEU_EINVAL, "null argument");
  if (window_bits > 16) return fail(HEU_EINVAL, "window_bits must be in [0, 16] (0 = automatic)");
  if (r_bits < 1 || r_bits > 32 * n_words) return fail(HEU_EINVAL, "r_bits must be in [1, bits(n)]");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  BnCtx bc;
  Bn n, n2, hs;
  bn_from_words(n, n_w, n_words);
  const int bits = BN_num_bits(n);
  const Ops* ops = heu_hip::ops_for_key_bits(bits);
  if (!ops) return fail(HEU_EUNSUPPORTED, "no kernels for a %d-bit modulus (2048 and 3072 are built)", bits);
  if (!BN_is_odd(n)) return fail(HEU_EINVAL, "n must be odd");
  BN_sqr(n2, n, bc.c);
  bn_from_words(hs, hs_w, 2 * (size_t)n_words);
  if (BN_cmp(hs, n2) >= 0) return fail(HEU_EINVAL, "h_s must be below n^2");
  auto key = std::make_unique<heu_key>();
  key->ctx = ctx;
  key->ops = ops;
  key->key_bits = bits;
  key->n_words = bits / 32;
  key->ct_words = ops->ct_words;
  key->prime_words = ops->p_words;
  key->nl_n = ops->n_L * ops->n_K;
  key->r_bits = r_bits;
  hipStream_t st = ctx->stream;
  size_t* tb = &key->device_bytes;
  int rc = upload_mod(n2, ops->n_L, ops->n_K, key->n2_m, key->n2_one, key->n2_r2, nullptr, &key->n2_mprime, st, tb, bc.c);
  if (rc) return rc;
  {
    // R^(k+1) mod n^2 for k < kRpowEntries: the Montgomery forms of R^k that
    // bring a chunk of k Z-space terms' Montgomery product into Montgomery form
    Bn R, r;
    BN_set_bit(R, heu_hip::kLimbBits * key->nl_n);
    BN_nnmod(r, R, n2, bc.c);
    std::vector<uint32_t> tab;
    tab.reserve((size_t)heu_hip::kRpowEntries * key->nl_n);
    for (int k = 0; k < heu_hip::kRpowEntries; ++k) {
      const std::vector<uint32_t> e = r29_interleaved(r, ops->n_L, ops->n_K);
      tab.insert(tab.end(), e.begin(), e.end());
      BN_mod_mul(r, r, R, n2, bc.c);
    }
    if ((rc = upload(key->n2_rpow, tab, st, tb))) return rc;
  }
  if ((rc = upload(key->n29, r29_plain(n, key->nl_n), st, tb))) return rc;
  key->n2_host = bn_to_words(n2, key->ct_words);
  if ((rc = upload(key->n2_words, key->n2_host, st, tb))) return rc;
  key->n_host = bn_to_words(n, 2 * ops->p_words);
  {
    Bn h;
    BN_rshift1(h, n);
    key->nhalf_host = bn_to_words(h, 2 * ops->p_words);
  }
  // fixed-base table; window 0 = the widest window whose table fits the
  // per-key budget (table_budget_bytes), at least the reference's 10
  if (window_bits == 0) {
    const size_t budget = table_budget_bytes();
    window_bits = 10;
    for (uint32_t w = 16; w > 10; --w) {
      const size_t bytes = ((size_t)((r_bits + w - 1) / w) << w) * (size_t)key->nl_n * 4;
      if (bytes <= budget) {
        window_bits = w;
        break;
      }
    }
  }
  key->hs_host = bn_to_words(hs, key->ct_words);
  if ((rc = build_table(key.get(), window_bits))) return rc;
  if (p_w && q_w) {
    Bn p, q, t, p2, q2;
    bn_from_words(p, p_w, prime_words);
    bn_from_words(q, q_w, prime_words);
    BN_mul(t, p, q, bc.c);
    if (BN_cmp(t, n) != 0) return fail(HEU_EKEY, "pk and sk are not paired: p*q != n");
    if (BN_num_bits(p) > 32 * ops->p_words || BN_num_bits(q) > 32 * ops->p_words)
      return fail(HEU_EINVAL, "prime wider than %d words", ops->p_words);
    BN_sqr(p2, p, bc.c);
    BN_sqr(q2, q, bc.c);
    for (int v = 0; v < heu_hip::kDecShapes; ++v) {  // constants of every decryption shape
      const heu_hip::DecShape& d = ops->dec[v];
      heu_key::Half &hp = key->p2[v], &hq = key->q2[v];
      if ((rc = upload_mod(p2, d.L, d.K, hp.m, hp.one, hp.r2, &hp.r3, &hp.mprime, st, tb, bc.c))) return rc;
      if ((rc = upload_mod(q2, d.L, d.K, hq.m, hq.one, hq.r2, &hq.r3, &hq.mprime, st, tb, bc.c))) return rc;
    }
    Bn pm1, qm1;
    BN_copy(pm1, p);
    BN_sub_word(pm1, 1);
    BN_copy(qm1, q);
    BN_sub_word(qm1, 1);
    const int dec_w = heu_hip::dec_window_bits(key->key_bits);
    auto sp = window_schedule(pm1, dec_w), sq = window_schedule(qm1, dec_w);
    key->nsched_p = (int)sp.size();
    key->nsched_q = (int)sq.size();
    if ((rc = upload(key->sched_p, sp, st, tb)) || (rc = upload(key->sched_q, sq, st, tb))) return rc;
    // CRT constants (secret_key.cc:19-43)
    const int PW = ops->p_words;
    Bn R, g, hp, hq, rem, pinv, qinv, pinvq, x;
    BN_set_bit(R, 32 * PW);
    BN_copy(g, n);
    BN_add_word(g, 1);
    auto h_of = [&](BIGNUM* hout, const BIGNUM* pr, const BIGNUM* pr2, const BIGNUM* prm1) {
      Bn u, d;
      BN_mod_exp(u, g, prm1, pr2, bc.c);
      BN_sub_word(u, 1);
      BN_div(u, d, u, pr, bc.c);
      BN_mod_inverse(hout, u, pr, bc.c);
    };
    h_of(hp, p, p2, pm1);
    h_of(hq, q, q2, qm1);
    BN_mod_inverse(pinv, p, R, bc.c);
    BN_mod_inverse(qinv, q, R, bc.c);
    BN_mod_inverse(pinvq, p, q, bc.c);
    auto up = [&](DevBuf& d, const BIGNUM* v, int words) { return upload(d, bn_to_words(v, words), st, tb); };
    if ((rc = up(key->c_p, p, PW)) || (rc = up(key->c_q, q, PW)) || (rc = up(key->c_pinv, pinv, PW)) ||
        (rc = up(key->c_qinv, qinv, PW)))
      return rc;
    BN_mod_mul(x, hp, R, p, bc.c);
    if ((rc = up(key->c_hpR, x, PW))) return rc;
    BN_mod_mul(x, hq, R, q, bc.c);
    if ((rc = up(key->c_hqR, x, PW))) return rc;
    BN_mod_mul(x, pinvq, R, q, bc.c);
    if ((rc = up(key->c_pinvqR, x, PW))) return rc;
    if ((rc = upload(key->c_n, key->n_host, st, tb)) || (rc = upload(key->c_nhalf, key->nhalf_host, st, tb))) return rc;
    key->c_pprime = (uint32_t)neg_inv_pow2(p, 32);
    key->c_qprime = (uint32_t)neg_inv_pow2(q, 32);
    key->has_sk = 1;
  }
  ctx->refs += 1;  // under ctx->mu: the key holds the context
  *out = key.release();
  return HEU_OK;
}

int heu_key_destroy(heu_key* key) {
  if (!key) return HEU_OK;
  heu_ctx* ctx = key->ctx;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();  // any stream (batches, servers) may still read the key
    delete key;
  }
  ctx_release(ctx);
  return HEU_OK;
}

int heu_key_info(const heu_key* key, uint32_t* key_bits, uint32_t* n_words, uint32_t* ct_words, uint32_t* has_sk) {
  if (!key) return fail(HEU_EINVAL, "null key");
  if (key_bits) *key_bits = (uint32_t)key->key_bits;
  if (n_words) *n_words = (uint32_t)key->n_words;
  if (ct_words) *ct_words = (uint32_t)key->ct_words;
  if (has_sk) *has_sk = (uint32_t)key->has_sk;
  return HEU_OK;
}

size_t heu_key_device_bytes(const heu_key* key) { return key ? key->device_bytes : 0; }

int heu_key_set_param(heu_key* key, int param, uint32_t value) {
  int rc = check_key(key);
  if (rc) return rc;
  heu_ctx* ctx = key->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIP_TRY(hipSetDevice(ctx->device));
  switch (param) {
    case HEU_PARAM_TABLE_WINDOW:
      if (value < 1 || value > 16) return fail(HEU_EINVAL, "table window must be 1..16 bits");
      return build_table(key, value);
    case HEU_PARAM_DEC_LANES:
      for (int v = 0; v < heu_hip::kDecShapes; ++v)
        if ((uint32_t)key->ops->dec[v].L == value) {
          HIP_TRY(hipDeviceSynchronize());
          key->dec_shape = v;
          return HEU_OK;
        }
      return fail(HEU_EINVAL, "no decryption shape with %u lanes for %d-bit keys", value, key->key_bits);
    default:
      return fail(HEU_EINVAL, "unknown parameter %d", param);
  }
}

int heu_key_get_param(const heu_key* key, int param, uint32_t* value) {
  if (!key || !value) return fail(HEU_EINVAL, "null argument");
  switch (param) {
    case HEU_PARAM_TABLE_WINDOW: *value = (uint32_t)key->unit; return HEU_OK;
    case HEU_PARAM_DEC_LANES: *value = (uint32_t)key->dshape().L; return HEU_OK;
    default: return fail(HEU_EINVAL, "unknown parameter %d", param);
  }
}

// Every key entry point: validate, serialize on the context, select its
// device, and drop any error another library left on this thread, so that the
// hipGetLastError() checks after launches report this call's own launches.
#define KEY_PROLOGUE(key)                          \
  int rc_ = check_key(key);                        \
  if (rc_) return rc_;                             \
  heu_ctx* ctx = (key)->ctx;                       \
  std::lock_guard<std::mutex> lk_(ctx->mu);        \
  HIP_TRY(hipSetDevice(ctx->device));              \
  (void)hipGetLastError();                         \
  const Ops* o = (key)->ops;                       \
  (void)o;

int heu_encrypt_i64(heu_key* key, const int64_t* m, const uint32_t* r, uint32_t r_words, size_t count, uint32_t* ct) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!m || !r || !ct || r_words == 0) return fail(HEU_EINVAL, "null argument");
  // |m| <= n/2 always holds for 64-bit m with n >= 2^66
  LAUNCH(ctx, "encrypt_i64", o->encrypt_i64(ctx->stream, key->pub(), m, r, (int)r_words, (long)count, ct, ctx->d_err));
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

int heu_encrypt(heu_key* key, const uint32_t* m, const uint32_t* m_host, uint32_t m_words, const uint32_t* r,
                uint32_t r_words, size_t count, uint32_t* ct) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!m || !m_host || !r || !ct || m_words == 0 || r_words == 0) return fail(HEU_EINVAL, "null argument");
  if (m_words > (uint32_t)key->ct_words) return fail(HEU_EINVAL, "m_words larger than a ciphertext");
  // encryptor.cc:38-40: |m| <= n/2
  BnCtx bc;
  Bn half, v;
  bn_from_words(half, key->nhalf_host.data(), key->nhalf_host.size());
  std::vector<uint32_t> mag(m_words);
  for (size_t i = 0; i < count; ++i) {
    const uint32_t* w = m_host + i * m_words;
    bool neg = (w[m_words - 1] >> 31) != 0;
    uint64_t c = 1;
    for (uint32_t j = 0; j < m_words; ++j) {
      uint32_t x = w[j];
      if (neg) {
        uint64_t s = (uint64_t)(uint32_t)~x + c;
        x = (uint32_t)s;
        c = s >> 32;
      }
      mag[j] = x;
    }
    bn_from_words(v, mag.data(), m_words);
    if (BN_cmp(v, half) > 0)
      return fail(HEU_ERANGE, "message number out of range at index %zu, max (abs) = n/2", i);
  }
  LAUNCH(ctx, "encrypt_big", o->encrypt_big(ctx->stream, key->pub(), m, (int)m_words, r, (int)r_words, (long)count, ct, ctx->d_err));
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

int heu_randomize(heu_key* key, const uint32_t* ct, const uint32_t* r, uint32_t r_words, size_t count,
                  uint32_t* out) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!r || !out || r_words == 0) return fail(HEU_EINVAL, "null argument");
  LAUNCH(ctx, "randomize", o->randomize(ctx->stream, key->pub(), ct, r, (int)r_words, (long)count, out, ctx->d_err));
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

int heu_decrypt(heu_key* key, const uint32_t* ct, size_t count, uint32_t* out, uint32_t out_words, uint8_t* overflow) {
  KEY_PROLOGUE(key);
  if (!key->has_sk) return fail(HEU_EKEY, "key has no secret part");
  if (count == 0) return HEU_OK;
  if (!ct || !out || out_words == 0) return fail(HEU_EINVAL, "null argument");
  const long slice = std::min<long>((long)count, kSliceElems);
  const size_t HWB = (size_t)key->ct_words / 2 * 4;
  Scratch sc(ctx);
  size_t halves = sc.add(2 * (size_t)slice * HWB);
  size_t tab = sc.add(2 * (size_t)slice * ((size_t)key->nl_p() * 4 << (heu_hip::dec_window_bits(key->key_bits) - 1)));
  int rc = sc.commit();
  if (rc) return rc;
  const DecArgs D = key->dec();
  const CombArgs C = key->comb();
  for (long s0 = 0; s0 < (long)count; s0 += slice) {
    const long n = std::min<long>(slice, (long)count - s0);
    LAUNCH(ctx, "dec_half", key->dshape().dec_half(ctx->stream, D, ct + s0 * key->ct_words, n, sc.at<uint32_t>(halves), sc.at<uint32_t>(tab)));
    LAUNCH(ctx, "dec_combine", o->combine(ctx->stream, C, sc.at<uint32_t>(halves), n, out + s0 * out_words, (int)out_words,
               overflow ? overflow + s0 : nullptr));
  }
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

int heu_add(heu_key* key, const uint32_t* a, const uint32_t* b, size_t count, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!a || !b || !out) return fail(HEU_EINVAL, "null argument");
  LAUNCH(ctx, "mulmod", o->mulmod(ctx->stream, key->mod_n2(), a, b, (long)count, out));
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

int heu_add_i64(heu_key* key, const uint32_t* ct, const int64_t* pt, size_t count, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!ct || !pt || !out) return fail(HEU_EINVAL, "null argument");
  LAUNCH(ctx, "add_pt_i64", o->add_pt_i64(ctx->stream, key->pub(), ct, pt, (long)count, out));
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

int heu_negate(heu_key* key, const uint32_t* a, size_t count, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!a || !out) return fail(HEU_EINVAL, "null argument");
  Scratch sc(ctx);
  size_t base = sc.add(inv_scratch_bytes(key, (long)count));
  int rc = sc.commit();
  if (rc) return rc;
  return batch_invert(key, a, nullptr, (long)count, out, base);
}

int heu_sub(heu_key* key, const uint32_t* a, const uint32_t* b, size_t count, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!a || !b || !out) return fail(HEU_EINVAL, "null argument");
  Scratch sc(ctx);
  size_t binv = sc.add(count * (size_t)key->ct_words * 4);
  size_t base = sc.add(inv_scratch_bytes(key, (long)count));
  int rc = sc.commit();
  if (rc) return rc;
  if ((rc = batch_invert(key, b, nullptr, (long)count, sc.at<uint32_t>(binv), base))) return rc;
  LAUNCH(ctx, "mulmod", o->mulmod(ctx->stream, key->mod_n2(), a, sc.at<uint32_t>(binv), (long)count, out));
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

int heu_mul_i64(heu_key* key, const uint32_t* ct, const int64_t* k, size_t count, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!ct || !k || !out) return fail(HEU_EINVAL, "null argument");
  const long slice = std::min<long>((long)count, kSliceElems);
  const size_t NLB = (size_t)key->nl_n * 4;
  Scratch sc(ctx);
  size_t tab = sc.add((size_t)slice * (NLB << heu_hip::kPowWindowBits));
  size_t neg = sc.add(count);
  size_t base = sc.add(inv_scratch_bytes(key, (long)count));
  int rc = sc.commit();
  if (rc) return rc;
  for (long s0 = 0; s0 < (long)count; s0 += slice) {
    const long n = std::min<long>(slice, (long)count - s0);
    LAUNCH(ctx, "pow_i64", o->pow_i64(ctx->stream, key->mod_n2(), ct + s0 * key->ct_words, k + s0, n, out + s0 * key->ct_words,
               sc.at<uint8_t>(neg) + s0, sc.at<uint32_t>(tab)));
  }
  HIP_TRY(hipGetLastError());
  // negative scalars: invert c^|k| (evaluator.cc:119, PowMod with a negative exponent)
  return batch_invert(key, out, sc.at<uint8_t>(neg), (long)count, out, base);
}

int heu_mul(heu_key* key, const uint32_t* ct, const uint32_t* k_mag, const uint8_t* k_neg, uint32_t k_words,
            size_t count, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!ct || !k_mag || !out || k_words == 0) return fail(HEU_EINVAL, "null argument");
  const long slice = std::min<long>((long)count, kSliceElems);
  const size_t NLB = (size_t)key->nl_n * 4;
  Scratch sc(ctx);
  size_t tab = sc.add((size_t)slice * (NLB << heu_hip::kPowWindowBits));
  size_t neg = sc.add(count);
  size_t base = sc.add(inv_scratch_bytes(key, (long)count));
  int rc = sc.commit();
  if (rc) return rc;
  for (long s0 = 0; s0 < (long)count; s0 += slice) {
    const long n = std::min<long>(slice, (long)count - s0);
    LAUNCH(ctx, "pow_words",
           o->pow_words(ctx->stream, key->mod_n2(), ct + s0 * key->ct_words, k_mag + s0 * k_words,
                        k_neg ? k_neg + s0 : nullptr, (int)k_words, n, out + s0 * key->ct_words,
                        sc.at<uint8_t>(neg) + s0, sc.at<uint32_t>(tab)));
  }
  HIP_TRY(hipGetLastError());
  return batch_invert(key, out, sc.at<uint8_t>(neg), (long)count, out, base);
}

int heu_sum(heu_key* key, const uint32_t* ct, const int64_t* offsets, size_t segments, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (!offsets || (segments && !out) || (!ct && segments && offsets[segments] > offsets[0])) return fail(HEU_EINVAL, "null argument");
  return segmented(key, ct, nullptr, offsets, segments, out);
}

int heu_dot_i64(heu_key* key, const uint32_t* ct, const int64_t* w, const int64_t* offsets, size_t segments,
                uint32_t* out) {
  KEY_PROLOGUE(key);
  if (!offsets || (segments && !out) || ((!ct || !w) && segments && offsets[segments] > offsets[0]))
    return fail(HEU_EINVAL, "null argument");
  return segmented(key, ct, w, offsets, segments, out);
}

int heu_sum_gather(heu_key* key, const uint32_t* ct, const int64_t* idx, const int64_t* offsets, size_t segments,
                   uint32_t* out) {
  KEY_PROLOGUE(key);
  if (!offsets || (segments && !out) || ((!ct || !idx) && segments && offsets[segments] > offsets[0]))
    return fail(HEU_EINVAL, "null argument");
  return segmented(key, ct, nullptr, offsets, segments, out, idx);
}

int heu_dot_gather(heu_key* key, const uint32_t* ct, const int64_t* idx, const int64_t* w, const int64_t* offsets,
                   size_t segments, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (!offsets || (segments && !out) || (segments && offsets[segments] > offsets[0] && (!ct || !w || !idx)))
    return fail(HEU_EINVAL, "null argument");
  return segmented(key, ct, w, offsets, segments, out, idx);
}

int heu_prefix_mul(heu_key* key, const uint32_t* ct, const int64_t* offsets, size_t segments, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (segments == 0) return HEU_OK;
  if (!offsets) return fail(HEU_EINVAL, "null argument");
  if (offsets[segments] == offsets[0]) return HEU_OK;  // no terms: nothing to write
  if (!ct || !out) return fail(HEU_EINVAL, "null argument");
  std::vector<Chunk> segs(segments);
  for (size_t s = 0; s < segments; ++s) {
    if (offsets[s + 1] < offsets[s]) return fail(HEU_EINVAL, "offsets must be non-decreasing");
    segs[s] = Chunk{offsets[s], offsets[s + 1] - offsets[s]};
  }
  Scratch sc(ctx);
  size_t off = sc.add(segments * sizeof(Chunk));
  int rc = sc.commit();
  if (rc) return rc;
  if ((rc = upload_meta(ctx, sc, off, segs.data(), segments * sizeof(Chunk)))) return rc;
  LAUNCH(ctx, "prefix_mul", o->prefix_mul(ctx->stream, key->mod_n2(), ct, sc.at<Chunk>(off), (long)segments, out));
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

static int mont_convert(heu_key* key, const uint32_t* ct, size_t count, uint32_t word_bits, uint32_t* out,
                        bool to_mont) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!ct || !out || word_bits == 0 || word_bits > 64) return fail(HEU_EINVAL, "bad argument");
  BnCtx bc;
  Bn n2, R;
  bn_from_words(n2, key->n2_host.data(), key->n2_host.size());
  const int bits = BN_num_bits(n2);
  const int rbits = (int)((bits + word_bits - 1) / word_bits * word_bits);
  BN_set_bit(R, rbits);
  BN_nnmod(R, R, n2, bc.c);
  if (!to_mont) BN_mod_inverse(R, R, n2, bc.c);
  Scratch sc(ctx);
  size_t cz = sc.add((size_t)key->ct_words * 4);
  int rc = sc.commit();
  if (rc) return rc;
  auto words = bn_to_words(R, key->ct_words);
  if ((rc = upload_meta(ctx, sc, cz, words.data(), words.size() * 4))) return rc;
  LAUNCH(ctx, "mul_const", o->mul_const(ctx->stream, key->mod_n2(), ct, sc.at<uint32_t>(cz), (long)count, out));
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

int heu_to_mont(heu_key* key, const uint32_t* ct, size_t count, uint32_t word_bits, uint32_t* out) {
  return mont_convert(key, ct, count, word_bits, out, true);
}

int heu_from_mont(heu_key* key, const uint32_t* ct, size_t count, uint32_t word_bits, uint32_t* out) {
  return mont_convert(key, ct, count, word_bits, out, false);
}

int heu_key_mform_words(const heu_key* key, uint32_t* words) {
  if (!key || !words) return fail(HEU_EINVAL, "null argument");
  *words = (uint32_t)key->nl_n;
  return HEU_OK;
}

int heu_to_mform(heu_key* key, const uint32_t* ct, size_t count, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!ct || !out) return fail(HEU_EINVAL, "null argument");
  LAUNCH(ctx, "to_mform", o->to_mform(ctx->stream, key->mod_n2(), ct, (long)count, out));
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

int heu_from_mform(heu_key* key, const uint32_t* mf, size_t count, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!mf || !out) return fail(HEU_EINVAL, "null argument");
  // x R -> x, canonical words: the Montgomery-input finish of a segment product
  LAUNCH(ctx, "from_mform", o->seg_finish(ctx->stream, key->mod_n2(), mf, (long)count, out));
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

int heu_add_mform(heu_key* key, const uint32_t* a, const uint32_t* b, size_t count, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (count == 0) return HEU_OK;
  if (!a || !b || !out) return fail(HEU_EINVAL, "null argument");
  LAUNCH(ctx, "mulmod_m", o->mulmod_m(ctx->stream, key->mod_n2(), a, b, (long)count, out));
  HIP_TRY(hipGetLastError());
  return HEU_OK;
}

int heu_sum_mform(heu_key* key, const uint32_t* mf, const int64_t* offsets, size_t segments, uint32_t* out) {
  KEY_PROLOGUE(key);
  if (!offsets || (segments && !out) || (!mf && segments && offsets[segments] > offsets[0]))
    return fail(HEU_EINVAL, "null argument");
  return segmented(key, mf, nullptr, offsets, segments, out, nullptr, nullptr, 0, true);
}

}  // extern "C"

namespace heu_engine {

int sum_fixed_scratch(heu_key* key, const uint32_t* ct, const int64_t* offsets, size_t segments, uint32_t* out,
                      char* buf, size_t bytes) {
  KEY_PROLOGUE(key);
  if (!offsets || (segments && !out) || (!ct && segments && offsets[segments] > offsets[0]))
    return fail(HEU_EINVAL, "null argument");
  return segmented(key, ct, nullptr, offsets, segments, out, nullptr, buf, bytes);
}

}  // namespace heu_engine
